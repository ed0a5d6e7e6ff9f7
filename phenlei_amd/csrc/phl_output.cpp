// phl_output.cpp -- field output and restart files of the residual + LU-SGS
// path (host side).
//
// Tecplot: the reference dumps node values of its visual variables as one
// FEPolyhedron zone per block (Post_WriteTecplot::StoreUnsFieldVisualData,
// API/PostProcess/src/Post_WriteTecplot.cpp:398-456; written by
// DumpToTecplotASCII::DumpFlowByFaceTopo, API/PostProcess/src/Post_Visual.cpp:
// 1801-1940).  Cell values reach the nodes by inverse-distance weighting over
// the cells around each node and the ghost cells of its non-symmetry boundary
// faces (CompNodeVar, API/Geometry/src/Geo_UnstructGrid.cpp:8527-8603).
//
// Restart: NSSolverUnstruct::DumpRestart (CFD/NSSolver/src/NSSolverUnstruct.cpp:
// 1179-1197) writes, per zone, int outnstep then q[m][nTotal] for the five
// primitives; TurbSolverUnstr::DumpRestart (CFD/TurbSolver/src/TurbSolverUnstr.
// cpp:435-456) writes its own block, int outnstep then nu~[nTotal].  A file
// here holds exactly one such block (steady flow) for one zone.
#include "../../include/phl.h"
#include "phl_mesh.h"

#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

namespace {

constexpr double kSmall = 1.0e-40;  // SMALL of the reference's node weights

// CompNodeVar with isVelocityForPostVisual false: for each variable row
// v[k][nTotal], the node values out[k][nNode].
void node_values(const phl::Mesh &m, int nVar, const std::vector<std::vector<double>> &v,
                 std::vector<std::vector<double>> &out)
{
    const int nN = m.nNode;
    std::vector<double> w(nN, 0.0);
    out.assign(nVar, std::vector<double>(nN, 0.0));
    auto add = [&](int node, int cell) {
        const double dx = m.x[node] - m.xcc[cell], dy = m.y[node] - m.ycc[cell], dz = m.z[node] - m.zcc[cell];
        const double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
        const double wt = 1.0 / dist;
        for (int k = 0; k < nVar; ++k) out[k][node] += v[k][cell] * wt;
        w[node] += wt;
    };
    for (int c = 0; c < m.nCell; ++c)
        for (int j = m.cellNodePtr[c]; j < m.cellNodePtr[c + 1]; ++j) add(m.cellNodes[j], c);
    for (int f = 0; f < m.nBFace; ++f) {
        if (m.bcType[f] == PHL_BC_SYMMETRY) continue;
        for (int j = m.faceNodePtr[f]; j < m.faceNodePtr[f + 1]; ++j) add(m.faceNodes[j], m.right[f]);
    }
    for (int i = 0; i < nN; ++i)
        for (int k = 0; k < nVar; ++k) out[k][i] = w[i] > kSmall ? out[k][i] / w[i] : 0.0;
}

// Fixed-width blocks of five numbers a line, as the reference's ASCII dump.
void write_block(FILE *fp, const double *a, int n)
{
    for (int i = 0; i < n; ++i) {
        std::fprintf(fp, "%20.12e", a[i]);
        if ((i + 1) % 5 == 0) std::fputc('\n', fp);
    }
    if (n % 5 != 0) std::fputc('\n', fp);
}

}  // namespace

extern "C" phl_status phl_write_tecplot(const phl_mesh *mh, const double *q, double mach, double gamma,
                                        const char *title, const char *path)
{
    if (!mh || !q || !path || mach <= 0.0 || gamma <= 1.0) {
        phl::set_last_error("phl_write_tecplot: bad argument");
        return PHL_ERR_ARG;
    }
    const phl::Mesh &m = mh->m;
    const int nT = m.nTotal();
    // visual variables 0-6 and 15 of the reference's default list:
    // density, u, v, w, pressure, temperature, Mach number, pressure coefficient
    static const char *kNames[] = {"density", "u", "v", "w", "pressure", "temperature", "mach", "cp"};
    const int nVar = 8;
    std::vector<std::vector<double>> v(nVar, std::vector<double>(nT));
    const double R = 1.0 / (gamma * mach * mach);  // PerfectGas.cpp:1401
    const double pInf = 1.0 / (gamma * mach * mach);
    for (int c = 0; c < nT; ++c) {
        const double r = q[c], u = q[(size_t)nT + c], vv = q[(size_t)2 * nT + c], w = q[(size_t)3 * nT + c],
                     p = q[(size_t)4 * nT + c];
        v[0][c] = r; v[1][c] = u; v[2][c] = vv; v[3][c] = w; v[4][c] = p;
        v[5][c] = p / (R * r);
        v[6][c] = std::sqrt((u * u + vv * vv + w * w) / (gamma * p / r));
        v[7][c] = 2.0 * (p - pInf);  // freestream density and speed are 1
    }
    std::vector<std::vector<double>> nv;
    node_values(m, nVar, v, nv);

    FILE *fp = std::fopen(path, "w");
    if (!fp) {
        phl::set_last_error(std::string("phl_write_tecplot: cannot open ") + path);
        return PHL_ERR_ARG;
    }
    std::fprintf(fp, "title = \"%s\"\n", title ? title : "phenlei_amd");
    std::fprintf(fp, "variables = \"x\" \"y\" \"z\"");
    for (int k = 0; k < nVar; ++k) std::fprintf(fp, " \"%s\"", kNames[k]);
    std::fprintf(fp, "\n");
    const int nFaceNodes = m.faceNodePtr[m.nFace];
    std::fprintf(fp, "zone T = \"Zone0 BLK\"\n");
    std::fprintf(fp, "ZoneType = FEPolyhedron\n");
    std::fprintf(fp, "Nodes    = %d\nFaces    = %d\nElements = %d\n", m.nNode, m.nFace, m.nCell);
    std::fprintf(fp, "TotalNumFaceNodes = %d\nNumConnectedBoundaryFaces = 0\nTotalNumBoundaryConnections = 0\n",
                 nFaceNodes);
    write_block(fp, m.x.data(), m.nNode);
    write_block(fp, m.y.data(), m.nNode);
    write_block(fp, m.z.data(), m.nNode);
    for (int k = 0; k < nVar; ++k) write_block(fp, nv[k].data(), m.nNode);
    for (int f = 0; f < m.nFace; ++f) {
        std::fprintf(fp, "%d ", m.faceNodePtr[f + 1] - m.faceNodePtr[f]);
        if ((f + 1) % 5 == 0) std::fputc('\n', fp);
    }
    if (m.nFace % 5 != 0) std::fputc('\n', fp);
    for (int f = 0; f < m.nFace; ++f) {
        for (int j = m.faceNodePtr[f]; j < m.faceNodePtr[f + 1]; ++j) std::fprintf(fp, "%d ", m.faceNodes[j] + 1);
        std::fputc('\n', fp);
    }
    // left and right elements, 1-based; a ghost (right of a boundary face) is 0
    for (int f = 0; f < m.nFace; ++f) {
        std::fprintf(fp, "%d ", m.left[f] + 1);
        if ((f + 1) % 5 == 0) std::fputc('\n', fp);
    }
    if (m.nFace % 5 != 0) std::fputc('\n', fp);
    for (int f = 0; f < m.nFace; ++f) {
        int r = m.right[f] + 1;
        if (r > m.nCell || r < 0) r = 0;
        std::fprintf(fp, "%d ", r);
        if ((f + 1) % 5 == 0) std::fputc('\n', fp);
    }
    if (m.nFace % 5 != 0) std::fputc('\n', fp);
    const bool ok = std::ferror(fp) == 0;
    if (std::fclose(fp) != 0 || !ok) {
        phl::set_last_error(std::string("phl_write_tecplot: write failed: ") + path);
        return PHL_ERR_ARG;
    }
    return PHL_OK;
}

extern "C" phl_status phl_write_restart(const char *path, int outnstep, int n_eq, int n_total, const double *q)
{
    if (!path || !q || n_total < 1 || n_eq < 1) {
        phl::set_last_error("phl_write_restart: bad argument");
        return PHL_ERR_ARG;
    }
    FILE *fp = std::fopen(path, "wb");
    if (!fp) {
        phl::set_last_error(std::string("phl_write_restart: cannot open ") + path);
        return PHL_ERR_ARG;
    }
    const size_t n = (size_t)n_eq * n_total;
    bool ok = std::fwrite(&outnstep, sizeof(int), 1, fp) == 1;
    ok = ok && std::fwrite(q, sizeof(double), n, fp) == n;
    ok = (std::fclose(fp) == 0) && ok;
    if (!ok) {
        phl::set_last_error(std::string("phl_write_restart: write failed: ") + path);
        return PHL_ERR_ARG;
    }
    return PHL_OK;
}

extern "C" phl_status phl_read_restart(const char *path, int n_eq, int n_total, int *outnstep, double *q)
{
    if (!path || !q || !outnstep || n_total < 1 || n_eq < 1) {
        phl::set_last_error("phl_read_restart: bad argument");
        return PHL_ERR_ARG;
    }
    FILE *fp = std::fopen(path, "rb");
    if (!fp) {
        phl::set_last_error(std::string("phl_read_restart: cannot open ") + path);
        return PHL_ERR_ARG;
    }
    std::fseek(fp, 0, SEEK_END);
    const long size = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    const size_t n = (size_t)n_eq * n_total;
    const long want = (long)sizeof(int) + (long)(sizeof(double) * n);
    if (size != want) {
        std::fclose(fp);
        phl::set_last_error("phl_read_restart: " + std::string(path) + " holds " + std::to_string(size) +
                            " bytes, expected " + std::to_string(want) + " for this zone");
        return PHL_ERR_ARG;
    }
    bool ok = std::fread(outnstep, sizeof(int), 1, fp) == 1;
    ok = ok && std::fread(q, sizeof(double), n, fp) == n;
    std::fclose(fp);
    if (!ok) {
        phl::set_last_error(std::string("phl_read_restart: short read: ") + path);
        return PHL_ERR_ARG;
    }
    return PHL_OK;
}
