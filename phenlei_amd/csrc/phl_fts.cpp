// phl_fts.cpp -- the reference's partitioned grid files (.fts) and wall
// distance files (.wdt), read without the HDF5 library.
//
// Both are HDF5 files written with default property lists
// (HDF5IO::CreateAndWriteData, API/IO/src/IO_HDF5File.cpp:103-140): version-0
// superblock, groups as symbol tables (v1 B-tree + local heap), and every
// dataset one contiguous block of little-endian int32, float64 or char.  This
// file reads exactly that subset of the format (HDF5 File Format
// Specification, sections II.A superblock, III.A.1 v1 B-trees, III.B symbol
// table nodes, III.D local heaps, IV.A.1 v1 object headers, IV.A.2 messages
// 0x0001 dataspace, 0x0003 datatype, 0x0008 layout, 0x0010 continuation,
// 0x0011 symbol table) and refuses anything else with an error.
//
// A .fts file holds nBlocks zones "Grid-<i>" (IO_HDF5Read::ReadUnstructuredGrid,
// API/PreProcess/src/Pre_HDF5File.cpp:1750-1797): iDimensions (nodes, faces,
// cells), GridCoordinates, FaceTopology (faces with a negative left cell are
// reversed and their cells swapped on reading, :1799-1848), CellTopology and
// InterfaceInfo (interface faces of BC code -1 and their partners in other
// zones).  The zones are joined into one zone here: cells zone by zone, nodes
// shared between zones merged (same coordinates), the non-interface boundary
// faces first (zone by zone, file order), then every zone's interior faces,
// then one interior face per interface pair, oriented from the lower zone's
// cell.  A 2-D grid (two nodes per face) stays 2-D: the metrics follow the
// reference's 2-D routines.
#include "../../include/phl.h"
#include "phl_mesh.h"

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

namespace {

// ------------------------------------------------------------ HDF5 subset
struct Dataset {
    std::vector<uint64_t> shape;
    int cls = -1;          // 0 fixed point, 1 floating point, 3 string
    int size = 0;          // bytes per element
    uint64_t addr = 0;     // file offset of the data
    uint64_t nbytes = 0;
};

class H5 {
public:
    bool open(const std::string &path, std::string &err)
    {
        FILE *fp = std::fopen(path.c_str(), "rb");
        if (!fp) return fail(err, "cannot open " + path);
        std::fseek(fp, 0, SEEK_END);
        const long n = std::ftell(fp);
        std::fseek(fp, 0, SEEK_SET);
        if (n < 96) { std::fclose(fp); return fail(err, path + " is too short for an HDF5 file"); }
        b_.resize((size_t)n);
        const bool ok = std::fread(b_.data(), 1, b_.size(), fp) == b_.size();
        std::fclose(fp);
        if (!ok) return fail(err, "short read of " + path);
        static const unsigned char sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
        if (std::memcmp(b_.data(), sig, 8) != 0) return fail(err, path + " is not an HDF5 file");
        const int sbv = b_[8];
        if (sbv != 0 && sbv != 1) return fail(err, "HDF5 superblock version " + std::to_string(sbv) + " is not supported");
        so_ = b_[13];
        sl_ = b_[14];
        if (so_ != 8 || sl_ != 8) return fail(err, "HDF5 offsets/lengths other than 8 bytes are not supported");
        size_t p = 24 + (sbv == 1 ? 4 : 0);
        base_ = u64(p);
        root_ = base_ + u64(p + 32 + 8);  // root symbol table entry: link name offset, object header address
        path_ = path;
        return true;
    }

    uint64_t root() const { return root_; }

    // Children (name, object header address) of a group.
    bool children(uint64_t hdr, std::vector<std::pair<std::string, uint64_t>> &out, std::string &err) const
    {
        std::vector<Msg> msgs;
        if (!messages(hdr, msgs, err)) return false;
        uint64_t btree = 0, heap = 0;
        bool found = false;
        for (const Msg &m : msgs)
            if (m.type == 0x11 && m.size >= 16) { btree = base_ + u64(m.data); heap = base_ + u64(m.data + 8); found = true; }
        if (!found) return fail(err, "object is not a symbol-table group");
        out.clear();
        return walk_btree(btree, heap, out, 0, err);
    }

    bool child(uint64_t grp, const std::string &name, uint64_t &hdr, std::string &err) const
    {
        std::vector<std::pair<std::string, uint64_t>> ch;
        if (!children(grp, ch, err)) return false;
        for (const auto &c : ch)
            if (c.first == name) { hdr = c.second; return true; }
        return fail(err, "no object '" + name + "' in " + path_);
    }

    bool dataset(uint64_t hdr, Dataset &d, std::string &err) const
    {
        std::vector<Msg> msgs;
        if (!messages(hdr, msgs, err)) return false;
        bool haveSpace = false, haveType = false, haveLayout = false;
        for (const Msg &m : msgs) {
            if (m.type == 0x0B) return fail(err, "filtered (compressed) datasets are not supported");
            if (m.type == 0x01) {
                if (!in(m.data, 8)) return fail(err, "bad dataspace message");
                const int ver = b_[m.data], rank = b_[m.data + 1];
                size_t p = m.data + (ver == 1 ? 8 : 4);
                if (ver != 1 && ver != 2) return fail(err, "dataspace version not supported");
                if (!in(p, (size_t)8 * rank)) return fail(err, "bad dataspace message");
                d.shape.clear();
                for (int i = 0; i < rank; ++i) d.shape.push_back(u64(p + 8 * i));
                haveSpace = true;
            } else if (m.type == 0x03) {
                if (!in(m.data, 8)) return fail(err, "bad datatype message");
                d.cls = b_[m.data] & 15;
                if ((d.cls == 0 || d.cls == 1) && (b_[m.data + 1] & 1)) return fail(err, "big-endian data is not supported");
                d.size = (int)u32(m.data + 4);
                haveType = true;
            } else if (m.type == 0x08) {
                if (!in(m.data, 2)) return fail(err, "bad layout message");
                const int ver = b_[m.data];
                if (ver == 3) {
                    const int lc = b_[m.data + 1];
                    if (lc == 1) {
                        if (!in(m.data + 2, 16)) return fail(err, "bad layout message");
                        d.addr = u64(m.data + 2);
                        d.nbytes = u64(m.data + 10);
                    } else if (lc == 0) {
                        if (!in(m.data + 2, 2)) return fail(err, "bad layout message");
                        d.nbytes = u16(m.data + 2);
                        d.addr = m.data + 4 - base_;
                    } else {
                        return fail(err, "chunked datasets are not supported");
                    }
                } else if (ver == 1 || ver == 2) {
                    if (!in(m.data, 8)) return fail(err, "bad layout message");
                    const int rank = b_[m.data + 1], lc = b_[m.data + 2];
                    if (lc != 1) return fail(err, "only contiguous layouts of version 1/2 are supported");
                    if (!in(m.data + 8, 8)) return fail(err, "bad layout message");
                    d.addr = u64(m.data + 8);
                    d.nbytes = 0;  // from the shape below
                    (void)rank;
                } else {
                    return fail(err, "layout version not supported");
                }
                haveLayout = true;
            }
        }
        if (!haveSpace || !haveType || !haveLayout) return fail(err, "object is not a dataset");
        uint64_t n = 1;
        for (uint64_t s : d.shape) {
            if (s > (1ull << 40) || (s && n > (1ull << 40) / s)) return fail(err, "dataset shape too large");
            n *= s;
        }
        const uint64_t want = n * (uint64_t)d.size;
        if (d.nbytes == 0) d.nbytes = want;
        if (d.nbytes < want) return fail(err, "dataset storage smaller than its shape");
        if (d.addr == ~0ull) return fail(err, "dataset without storage");
        d.addr += base_;
        if (!in(d.addr, want)) return fail(err, "dataset runs past the end of " + path_);
        d.nbytes = want;
        return true;
    }

    bool read_ints(uint64_t grp, const char *name, std::vector<int> &v, std::string &err) const
    {
        uint64_t h = 0;
        Dataset d;
        if (!child(grp, name, h, err) || !dataset(h, d, err)) return false;
        if (d.cls != 0 || d.size != 4) return fail(err, std::string(name) + " is not an int32 dataset");
        v.resize(d.nbytes / 4);
        if (!v.empty()) std::memcpy(v.data(), b_.data() + d.addr, d.nbytes);
        return true;
    }

    bool read_doubles(uint64_t grp, const char *name, std::vector<double> &v, std::string &err) const
    {
        uint64_t h = 0;
        Dataset d;
        if (!child(grp, name, h, err) || !dataset(h, d, err)) return false;
        return doubles(d, v, name, err);
    }

    bool doubles(const Dataset &d, std::vector<double> &v, const std::string &name, std::string &err) const
    {
        if (d.cls != 1 || d.size != 8) return fail(err, name + " is not a float64 dataset");
        v.resize(d.nbytes / 8);
        if (!v.empty()) std::memcpy(v.data(), b_.data() + d.addr, d.nbytes);
        return true;
    }

    bool read_int(uint64_t grp, const char *name, int &out, std::string &err) const
    {
        std::vector<int> v;
        if (!read_ints(grp, name, v, err)) return false;
        if (v.size() != 1) return fail(err, std::string(name) + " is not a scalar");
        out = v[0];
        return true;
    }

private:
    struct Msg {
        int type;
        size_t data, size;
    };

    static bool fail(std::string &err, const std::string &m)
    {
        err = m;
        return false;
    }
    bool in(size_t p, size_t n) const { return p <= b_.size() && n <= b_.size() - p; }
    uint64_t u64(size_t p) const { uint64_t v = 0; if (in(p, 8)) std::memcpy(&v, b_.data() + p, 8); return v; }
    uint32_t u32(size_t p) const { uint32_t v = 0; if (in(p, 4)) std::memcpy(&v, b_.data() + p, 4); return v; }
    uint16_t u16(size_t p) const { uint16_t v = 0; if (in(p, 2)) std::memcpy(&v, b_.data() + p, 2); return v; }

    // Messages of a version-1 object header, following continuation blocks.
    bool messages(uint64_t hdr, std::vector<Msg> &out, std::string &err) const
    {
        if (!in(hdr, 16)) return fail(err, "object header outside the file");
        if (b_[hdr] != 1) return fail(err, "object header version " + std::to_string(b_[hdr]) + " is not supported");
        const int nmsg = u16(hdr + 2);
        std::vector<std::pair<uint64_t, uint64_t>> blocks{{hdr + 16, u32(hdr + 8)}};
        out.clear();
        for (size_t k = 0; k < blocks.size() && (int)out.size() < nmsg; ++k) {
            if (k > 64) return fail(err, "too many object header continuation blocks");
            const uint64_t start = blocks[k].first, len = blocks[k].second;
            if (!in(start, len)) return fail(err, "object header block outside the file");
            uint64_t p = start;
            while (p + 8 <= start + len && (int)out.size() < nmsg) {
                const int type = u16(p), size = u16(p + 2);
                const uint64_t data = p + 8;
                if (data + size > start + len) return fail(err, "object header message runs past its block");
                if (type == 0x10) {
                    if (size < 16) return fail(err, "bad continuation message");
                    blocks.push_back({base_ + u64(data), u64(data + 8)});
                }
                out.push_back({type, (size_t)data, (size_t)size});
                p = data + size;
            }
        }
        return true;
    }

    bool heap_name(uint64_t heap, uint64_t off, std::string &name, std::string &err) const
    {
        if (!in(heap, 32) || std::memcmp(b_.data() + heap, "HEAP", 4) != 0) return fail(err, "bad local heap");
        const uint64_t dsize = u64(heap + 8), daddr = base_ + u64(heap + 24);
        if (off >= dsize || !in(daddr, dsize)) return fail(err, "name outside the local heap");
        const char *s = reinterpret_cast<const char *>(b_.data() + daddr + off);
        name.assign(s, strnlen(s, dsize - off));
        return true;
    }

    bool walk_btree(uint64_t node, uint64_t heap, std::vector<std::pair<std::string, uint64_t>> &out, int depth,
                    std::string &err) const
    {
        if (depth > 16) return fail(err, "B-tree too deep");
        if (!in(node, 24) || std::memcmp(b_.data() + node, "TREE", 4) != 0) return fail(err, "bad group B-tree node");
        if (b_[node + 4] != 0) return fail(err, "B-tree is not a group tree");
        const int level = b_[node + 5], used = u16(node + 6);
        size_t p = node + 24;
        if (!in(p, (size_t)16 * used + 8)) return fail(err, "B-tree node outside the file");
        for (int i = 0; i < used; ++i) {
            const uint64_t child = base_ + u64(p + 8);
            p += 16;
            if (level > 0) {
                if (!walk_btree(child, heap, out, depth + 1, err)) return false;
                continue;
            }
            if (!in(child, 8) || std::memcmp(b_.data() + child, "SNOD", 4) != 0) return fail(err, "bad symbol table node");
            const int nsym = u16(child + 6);
            if (!in(child + 8, (size_t)40 * nsym)) return fail(err, "symbol table node outside the file");
            for (int s = 0; s < nsym; ++s) {
                const size_t e = child + 8 + 40 * (size_t)s;
                std::string name;
                if (!heap_name(heap, u64(e), name, err)) return false;
                out.push_back({name, base_ + u64(e + 8)});
            }
        }
        return true;
    }

    std::vector<unsigned char> b_;
    int so_ = 8, sl_ = 8;
    uint64_t base_ = 0, root_ = 0;
    std::string path_;
};

// ------------------------------------------------------------ zones
struct Zone {
    int nNode = 0, nFace = 0, nCell = 0, nBFace = 0;
    std::vector<double> x, y, z;
    std::vector<int> faceNodeCount, faceNodes, left, right, bc;
    std::vector<int> cellNodeCount, cellNodes;
    std::vector<int> ifZone, ifPartner, ifFace;  // InterfaceInfo
    std::vector<int> facePtr;
};

constexpr int kBcInterface = -1;  // PHENGLEI::INTERFACE

bool read_zone(const H5 &h, uint64_t zg, Zone &z, std::string &err)
{
    std::vector<int> dims;
    if (!h.read_ints(zg, "iDimensions", dims, err)) return false;
    if (dims.size() != 3 || dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0) { err = "bad iDimensions"; return false; }
    z.nNode = dims[0]; z.nFace = dims[1]; z.nCell = dims[2];
    uint64_t g = 0;
    if (!h.child(zg, "GridCoordinates", g, err) || !h.read_doubles(g, "CoordinateX", z.x, err) ||
        !h.read_doubles(g, "CoordinateY", z.y, err) || !h.read_doubles(g, "CoordinateZ", z.z, err))
        return false;
    if ((int)z.x.size() != z.nNode || (int)z.y.size() != z.nNode || (int)z.z.size() != z.nNode) {
        err = "coordinate arrays do not match iDimensions";
        return false;
    }
    if (!h.child(zg, "FaceTopology", g, err) || !h.read_ints(g, "nodeNumberOfEachFace", z.faceNodeCount, err) ||
        !h.read_ints(g, "face2Node", z.faceNodes, err) || !h.read_ints(g, "leftCellOfFace", z.left, err) ||
        !h.read_ints(g, "rightCellOfFace", z.right, err) || !h.read_int(g, "nBoundFace", z.nBFace, err) ||
        !h.read_ints(g, "BcType", z.bc, err))
        return false;
    if ((int)z.faceNodeCount.size() != z.nFace || (int)z.left.size() != z.nFace || (int)z.right.size() != z.nFace ||
        z.nBFace < 0 || z.nBFace > z.nFace || (int)z.bc.size() != z.nBFace) {
        err = "face arrays do not match iDimensions";
        return false;
    }
    z.facePtr.assign(z.nFace + 1, 0);
    for (int f = 0; f < z.nFace; ++f) {
        if (z.faceNodeCount[f] < 2) { err = "face with fewer than two nodes"; return false; }
        z.facePtr[f + 1] = z.facePtr[f] + z.faceNodeCount[f];
    }
    if (z.facePtr[z.nFace] != (int)z.faceNodes.size()) { err = "face2Node size does not match"; return false; }
    for (int n : z.faceNodes)
        if (n < 0 || n >= z.nNode) { err = "face node index out of range"; return false; }
    // IO_HDF5Read::ReadFaceTopology: a face whose left cell is negative is
    // reversed and its cells swapped.
    for (int f = 0; f < z.nFace; ++f)
        if (z.left[f] < 0) {
            std::reverse(z.faceNodes.begin() + z.facePtr[f], z.faceNodes.begin() + z.facePtr[f + 1]);
            std::swap(z.left[f], z.right[f]);
        }
    for (int f = 0; f < z.nFace; ++f) {
        const bool ok = z.left[f] >= 0 && z.left[f] < z.nCell &&
                        (f < z.nBFace || (z.right[f] >= 0 && z.right[f] < z.nCell && z.right[f] != z.left[f]));
        if (!ok) { err = "face " + std::to_string(f) + " has bad cells"; return false; }
    }
    if (!h.child(zg, "CellTopology", g, err) || !h.read_ints(g, "nodeNumberOfEachCell", z.cellNodeCount, err) ||
        !h.read_ints(g, "cell2Node", z.cellNodes, err))
        return false;
    if ((int)z.cellNodeCount.size() != z.nCell) { err = "cell arrays do not match iDimensions"; return false; }
    long sum = 0;
    for (int k : z.cellNodeCount) sum += k;
    if (sum != (long)z.cellNodes.size()) { err = "cell2Node size does not match"; return false; }
    for (int n : z.cellNodes)
        if (n < 0 || n >= z.nNode) { err = "cell node index out of range"; return false; }
    int nIFace = 0;
    if (h.child(zg, "InterfaceInfo", g, err) && h.read_int(g, "nIFace", nIFace, err) && nIFace > 0) {
        if (!h.read_ints(g, "interFace2ZoneID", z.ifZone, err) ||
            !h.read_ints(g, "interFace2InterFaceID", z.ifPartner, err) ||
            !h.read_ints(g, "interFace2BoundaryFace", z.ifFace, err))
            return false;
        if ((int)z.ifZone.size() != nIFace || (int)z.ifPartner.size() != nIFace || (int)z.ifFace.size() != nIFace) {
            err = "interface arrays do not match nIFace";
            return false;
        }
    }
    err.clear();
    for (int i = 0; i < (int)z.ifFace.size(); ++i)
        if (z.ifFace[i] < 0 || z.ifFace[i] >= z.nBFace || z.bc[z.ifFace[i]] != kBcInterface) {
            err = "interface " + std::to_string(i) + " does not name an interface boundary face";
            return false;
        }
    return true;
}

bool join_zones(std::vector<Zone> &zs, phl::Mesh &m, std::string &err)
{
    const int nZ = (int)zs.size();
    std::vector<int> cellOff(nZ + 1, 0);
    for (int k = 0; k < nZ; ++k) cellOff[k + 1] = cellOff[k] + zs[k].nCell;
    const int dimFaces = zs[0].faceNodeCount[0];
    bool twoD = true;
    for (const Zone &z : zs)
        for (int c : z.faceNodeCount) twoD = twoD && c == 2;
    (void)dimFaces;

    // nodes: one per distinct coordinate triple across zones
    std::map<std::tuple<uint64_t, uint64_t, uint64_t>, int> key;
    std::vector<std::vector<int>> nodeMap(nZ);
    m.x.clear(); m.y.clear(); m.z.clear();
    for (int k = 0; k < nZ; ++k) {
        const Zone &z = zs[k];
        nodeMap[k].resize(z.nNode);
        for (int n = 0; n < z.nNode; ++n) {
            uint64_t bx, by, bz;
            std::memcpy(&bx, &z.x[n], 8); std::memcpy(&by, &z.y[n], 8); std::memcpy(&bz, &z.z[n], 8);
            auto ins = key.emplace(std::make_tuple(bx, by, bz), (int)m.x.size());
            if (ins.second) { m.x.push_back(z.x[n]); m.y.push_back(z.y[n]); m.z.push_back(z.z[n]); }
            nodeMap[k][n] = ins.first->second;
        }
    }
    m.nNode = (int)m.x.size();
    m.nCell = cellOff[nZ];

    struct F { int zone, face, left, right; bool flip; };
    std::vector<F> bnd, inner;
    for (int k = 0; k < nZ; ++k)
        for (int f = 0; f < zs[k].nBFace; ++f)
            if (zs[k].bc[f] != kBcInterface) bnd.push_back({k, f, cellOff[k] + zs[k].left[f], -1, false});
    for (int k = 0; k < nZ; ++k)
        for (int f = zs[k].nBFace; f < zs[k].nFace; ++f)
            inner.push_back({k, f, cellOff[k] + zs[k].left[f], cellOff[k] + zs[k].right[f], false});
    std::vector<int> used;
    for (int k = 0; k < nZ; ++k) {
        const Zone &z = zs[k];
        int nIfBc = 0;
        for (int f = 0; f < z.nBFace; ++f) nIfBc += z.bc[f] == kBcInterface;
        if (nIfBc != (int)z.ifFace.size()) {
            err = "zone " + std::to_string(k) + ": interface faces and InterfaceInfo disagree";
            return false;
        }
        for (int i = 0; i < (int)z.ifFace.size(); ++i) {
            const int nz = z.ifZone[i], ni = z.ifPartner[i];
            if (nz < 0 || nz >= nZ || ni < 0 || ni >= (int)zs[nz].ifFace.size() || zs[nz].ifZone[ni] != k ||
                zs[nz].ifPartner[ni] != i) {
                err = "zone " + std::to_string(k) + " interface " + std::to_string(i) + " has no matching partner";
                return false;
            }
            if (nz < k || (nz == k && ni < i)) continue;  // the pair is taken from its lower side
            const int f = z.ifFace[i], g = zs[nz].ifFace[ni];
            // same node set, opposite orientation
            std::vector<int> a, b;
            for (int j = z.facePtr[f]; j < z.facePtr[f + 1]; ++j) a.push_back(nodeMap[k][z.faceNodes[j]]);
            for (int j = zs[nz].facePtr[g]; j < zs[nz].facePtr[g + 1]; ++j) b.push_back(nodeMap[nz][zs[nz].faceNodes[j]]);
            std::vector<int> sa = a, sb = b;
            std::sort(sa.begin(), sa.end());
            std::sort(sb.begin(), sb.end());
            if (sa != sb) {
                err = "zone " + std::to_string(k) + " interface " + std::to_string(i) + ": partner face has other nodes";
                return false;
            }
            inner.push_back({k, f, cellOff[k] + z.left[f], cellOff[nz] + zs[nz].left[g], false});
        }
    }
    m.nBFace = (int)bnd.size();
    m.nFace = m.nBFace + (int)inner.size();
    m.faceNodePtr.assign(1, 0);
    m.faceNodes.clear();
    m.left.clear(); m.right.clear(); m.bcType.clear();
    auto put = [&](const F &fc, int right) {
        const Zone &z = zs[fc.zone];
        for (int j = z.facePtr[fc.face]; j < z.facePtr[fc.face + 1]; ++j) m.faceNodes.push_back(nodeMap[fc.zone][z.faceNodes[j]]);
        m.faceNodePtr.push_back((int)m.faceNodes.size());
        m.left.push_back(fc.left);
        m.right.push_back(right);
    };
    for (int i = 0; i < m.nBFace; ++i) {
        put(bnd[i], m.nCell + i);
        m.bcType.push_back(zs[bnd[i].zone].bc[bnd[i].face]);
    }
    for (const F &fc : inner) put(fc, fc.right);
    m.cellNodePtr.assign(1, 0);
    m.cellNodes.clear();
    for (int k = 0; k < nZ; ++k) {
        const Zone &z = zs[k];
        int p = 0;
        for (int c = 0; c < z.nCell; ++c) {
            for (int j = 0; j < z.cellNodeCount[c]; ++j) m.cellNodes.push_back(nodeMap[k][z.cellNodes[p + j]]);
            p += z.cellNodeCount[c];
            m.cellNodePtr.push_back((int)m.cellNodes.size());
        }
    }
    m.dim = twoD ? 2 : 3;
    m.computeMetrics();
    for (int c = 0; c < m.nCell; ++c)
        if (!(m.vol[c] > 0.0)) {
            err = "cell " + std::to_string(c) + " has a non-positive volume (face orientation)";
            return false;
        }
    return true;
}

bool zone_names(const H5 &h, std::vector<uint64_t> &groups, std::string &err)
{
    uint64_t info = 0;
    int nBlocks = 0;
    if (!h.child(h.root(), "Information", info, err) || !h.read_int(info, "nBlocks", nBlocks, err)) return false;
    if (nBlocks < 1) { err = "no zones in the file"; return false; }
    groups.clear();
    for (int i = 0; i < nBlocks; ++i) {
        uint64_t g = 0;
        if (!h.child(h.root(), "Grid-" + std::to_string(i), g, err)) return false;
        groups.push_back(g);
    }
    return true;
}

}  // namespace

extern "C" phl_status phl_read_fts(const char *path, phl_mesh **out, int *n_zones)
{
    if (!path || !out) {
        phl::set_last_error("phl_read_fts: bad argument");
        return PHL_ERR_ARG;
    }
    std::string err;
    H5 h;
    std::vector<uint64_t> groups;
    if (!h.open(path, err) || !zone_names(h, groups, err)) {
        phl::set_last_error("phl_read_fts: " + err);
        return PHL_ERR_ARG;
    }
    std::vector<Zone> zs(groups.size());
    for (size_t i = 0; i < groups.size(); ++i)
        if (!read_zone(h, groups[i], zs[i], err)) {
            phl::set_last_error("phl_read_fts: Grid-" + std::to_string(i) + ": " + err);
            return PHL_ERR_ARG;
        }
    phl_mesh *pm = new (std::nothrow) phl_mesh;
    if (!pm) {
        phl::set_last_error("phl_read_fts: out of memory");
        return PHL_ERR_ALLOC;
    }
    if (!join_zones(zs, pm->m, err)) {
        delete pm;
        phl::set_last_error("phl_read_fts: " + err);
        return PHL_ERR_ARG;
    }
    if (n_zones) *n_zones = (int)zs.size();
    *out = pm;
    return PHL_OK;
}

extern "C" phl_status phl_read_wdt(const char *path, int n_cell, double *wdist)
{
    if (!path || !wdist || n_cell < 1) {
        phl::set_last_error("phl_read_wdt: bad argument");
        return PHL_ERR_ARG;
    }
    std::string err;
    H5 h;
    std::vector<std::pair<std::string, uint64_t>> ch;
    if (!h.open(path, err) || !h.children(h.root(), ch, err)) {
        phl::set_last_error("phl_read_wdt: " + err);
        return PHL_ERR_ARG;
    }
    // Pre_WalldistCompute::DumpWallDist writes one dataset "Grid-<i>" per zone
    std::vector<double> all;
    for (int i = 0;; ++i) {
        const std::string name = "Grid-" + std::to_string(i);
        auto it = std::find_if(ch.begin(), ch.end(), [&](const std::pair<std::string, uint64_t> &c) { return c.first == name; });
        if (it == ch.end()) break;
        Dataset d;
        std::vector<double> v;
        if (!h.dataset(it->second, d, err) || !h.doubles(d, v, name, err)) {
            phl::set_last_error("phl_read_wdt: " + err);
            return PHL_ERR_ARG;
        }
        all.insert(all.end(), v.begin(), v.end());
    }
    if ((int)all.size() != n_cell) {
        phl::set_last_error("phl_read_wdt: " + std::string(path) + " holds " + std::to_string(all.size()) +
                            " cell distances, expected " + std::to_string(n_cell));
        return PHL_ERR_ARG;
    }
    std::memcpy(wdist, all.data(), sizeof(double) * all.size());
    return PHL_OK;
}
