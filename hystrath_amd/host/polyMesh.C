// constant/polyMesh (points, faces, owner, neighbour, boundary) read and
// recognised as the engine's block.  The reference reads any polyMesh through
// OpenFOAM (polyMesh.C, fvMesh); this engine tracks on one Cartesian block, so
// a polyMesh is accepted when it is one and rejected with the reason otherwise.
#include <sys/stat.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <set>

#include "blockMesh.H"
#include "foamText.H"

namespace hs {

namespace {

const char* const kSide[6] = { "xmin", "xmax", "ymin", "ymax", "zmin", "zmax" };
const char* const kFiles[5] = { "points", "faces", "owner", "neighbour", "boundary" };

bool isFile(const std::string& p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

std::string meshDir(const std::string& caseDir) { return caseDir + "/constant/polyMesh"; }

std::vector<Vec3> readPoints(const std::string& path) {
    FoamText t(path);
    std::vector<Vec3> v;
    readList<Vec3>(t, vec3, [&](long long k, const Vec3& p) {
        if ((size_t)k >= v.size()) v.resize(k + 1);
        v[k] = p;
    });
    return v;
}

// faceList "k(p0 .. pk-1)"; every face must be a quad
std::vector<std::array<long long, 4>> readFaces(const std::string& path) {
    FoamText t(path);
    std::vector<std::array<long long, 4>> v;
    readList<std::array<long long, 4>>(
        t,
        [](FoamText& u) {
            const long long k = u.label();
            if (k != 4) u.fail("a face of " + std::to_string(k) + " points; only hex cells are read");
            std::array<long long, 4> q;
            u.expect('(');
            for (auto& x : q) x = u.label();
            u.expect(')');
            return q;
        },
        [&](long long k, const std::array<long long, 4>& q) {
            if ((size_t)k >= v.size()) v.resize(k + 1);
            v[k] = q;
        });
    return v;
}

std::vector<long long> readLabels(const std::string& path) {
    FoamText t(path);
    return labelList(t);
}

struct Patch {
    std::string name, type;
    long long nFaces = -1, startFace = -1;
};

// polyBoundaryMesh "N ( name { type t; nFaces n; startFace s; ... } ... )"
std::vector<Patch> readBoundary(const std::string& path) {
    FoamText t(path);
    const long long n = t.label();
    t.expect('(');
    std::vector<Patch> out;
    for (long long k = 0; k < n; ++k) {
        Patch p;
        p.name = t.word();
        t.expect('{');
        while (!t.peek('}')) {
            const std::string key = t.word();
            const std::string val = t.value();
            try {
                if (key == "type") p.type = val;
                else if (key == "nFaces") p.nFaces = std::stoll(val);
                else if (key == "startFace") p.startFace = std::stoll(val);
            } catch (const std::exception&) {
                t.fail("patch " + p.name + ": " + key + " is not a label");
            }
        }
        t.expect('}');
        if (p.type.empty() || p.nFaces < 0 || p.startFace < 0)
            t.fail("patch " + p.name + " needs type, nFaces and startFace");
        out.push_back(p);
    }
    t.expect(')');
    return out;
}

// Number of lattice planes that the cell centres x lie on.  Centres are
// grouped into planes by a gap wider than 1e-7 of their span (uniform
// lattices of up to ten million cells a side are resolved); the gaps between
// planes must agree to a relative 1e-9 (OpenFOAM writes points with about
// 1e-10 relative precision).
int latticePlanes(std::vector<double> x, const std::string& what) {
    std::sort(x.begin(), x.end());
    const double span = x.back() - x.front();
    if (span <= 1e-12 * std::max({ std::fabs(x.front()), std::fabs(x.back()), 1e-300 })) return 1;
    std::vector<double> planes;
    size_t b = 0;
    for (size_t i = 1; i <= x.size(); ++i) {
        if (i < x.size() && x[i] - x[i - 1] <= 1e-7 * span) continue;
        if (x[i - 1] - x[b] > 1e-7 * span) throw FatalError(what + ": cell centres do not lie on planes");
        double s = 0;
        for (size_t k = b; k < i; ++k) s += x[k];
        planes.push_back(s / (double)(i - b));
        b = i;
    }
    const double mean = (planes.back() - planes.front()) / (double)(planes.size() - 1);
    for (size_t k = 1; k < planes.size(); ++k)
        if (std::fabs(planes[k] - planes[k - 1] - mean) > 1e-9 * std::fabs(mean) + 1e-12 * span)
            throw FatalError(what + ": cells are not evenly spaced (simpleGrading other than 1)");
    return (int)planes.size();
}

}  // namespace

bool hasPolyMesh(const std::string& caseDir) {
    for (const char* f : kFiles)
        if (isFile(meshDir(caseDir) + "/" + f) || isFile(meshDir(caseDir) + "/" + f + ".gz")) return true;
    return false;
}

BlockMesh readPolyMesh(const std::string& caseDir) {
    const std::string d = meshDir(caseDir);
    for (const char* f : kFiles)
        if (!isFile(d + "/" + f))
            throw FatalError(d + ": no " + f + (isFile(d + "/" + f + ".gz") ? " (compressed files are not read)" : ""));
    const auto pts = readPoints(d + "/points");
    const auto faces = readFaces(d + "/faces");
    const auto owner = readLabels(d + "/owner");
    const auto neigh = readLabels(d + "/neighbour");
    const auto patches = readBoundary(d + "/boundary");
    const size_t nf = faces.size();
    if (owner.size() != nf)
        throw FatalError(d + ": " + std::to_string(owner.size()) + " owners for " + std::to_string(nf) + " faces");
    if (neigh.size() > nf || nf == 0) throw FatalError(d + ": more neighbours than faces");
    long long nc = 0;
    for (long long c : owner) nc = std::max(nc, c + 1);
    for (long long c : owner)
        if (c < 0) throw FatalError(d + ": a negative owner");
    for (long long c : neigh)
        if (c < 0 || c >= nc) throw FatalError(d + ": a neighbour outside the cells");
    for (const auto& q : faces)
        for (long long p : q)
            if (p < 0 || p >= (long long)pts.size()) throw FatalError(d + ": a face point outside the points");

    // face centres, then cell centres as the mean of their six face centres
    // (exact for a box)
    std::vector<std::array<double, 3>> fc(nf), cc(nc, { 0.0, 0.0, 0.0 });
    for (size_t f = 0; f < nf; ++f) {
        const auto& q = faces[f];
        const Vec3 &a = pts[q[0]], &b = pts[q[1]], &c = pts[q[2]], &e = pts[q[3]];
        fc[f] = { (a.x + b.x + c.x + e.x) * 0.25, (a.y + b.y + c.y + e.y) * 0.25, (a.z + b.z + c.z + e.z) * 0.25 };
    }
    std::vector<int> cnt(nc, 0);
    auto add = [&](long long c, size_t f) {
        ++cnt[c];
        for (int a = 0; a < 3; ++a) cc[c][a] += fc[f][a];
    };
    for (size_t f = 0; f < nf; ++f) add(owner[f], f);
    for (size_t f = 0; f < neigh.size(); ++f) add(neigh[f], f);
    std::set<int> odd;
    for (int k : cnt)
        if (k != 6) odd.insert(k);
    if (!odd.empty()) {
        std::string s;
        for (int k : odd) s += (s.empty() ? "" : ", ") + std::to_string(k);
        throw FatalError(d + ": cells with " + s + " faces; only hex cells are read");
    }
    for (auto& c : cc)
        for (double& x : c) x /= 6.0;

    BlockMesh m;
    double spacing[3];
    for (int a = 0; a < 3; ++a) {
        std::vector<double> x(nc);
        for (long long c = 0; c < nc; ++c) x[c] = cc[c][a];
        m.n[a] = latticePlanes(std::move(x), d + " (" + "xyz"[a] + ")");
        auto coord = [a](const Vec3& p) { return a == 0 ? p.x : a == 1 ? p.y : p.z; };
        double lo = coord(pts[0]), hi = lo;
        for (const Vec3& p : pts) {
            lo = std::min(lo, coord(p));
            hi = std::max(hi, coord(p));
        }
        m.lo[a] = lo;     // exact, from the points
        m.hi[a] = hi;
        spacing[a] = (hi - lo) / m.n[a];
    }
    if ((long long)m.n[0] * m.n[1] * m.n[2] != nc)
        throw FatalError(d + ": " + std::to_string(nc) + " cells do not fill a " + std::to_string(m.n[0]) + "x" +
                         std::to_string(m.n[1]) + "x" + std::to_string(m.n[2]) + " block");
    for (long long c = 0; c < nc; ++c) {
        long long ijk[3];
        for (int a = 0; a < 3; ++a) ijk[a] = (long long)std::floor((cc[c][a] - m.lo[a]) / spacing[a]);
        if (ijk[0] + m.n[0] * (ijk[1] + m.n[1] * ijk[2]) != c)
            throw FatalError(d + ": the cell numbering is not blockMesh's i + nx*(j + ny*k)");
    }

    // each patch on whole sides of the block, each side held by one patch
    for (const Patch& p : patches) {
        m.patchNames.push_back(p.name);
        if (p.nFaces == 0) continue;
        if (p.startFace < (long long)neigh.size() || p.startFace + p.nFaces > (long long)nf)
            throw FatalError(d + ": patch " + p.name + " holds faces outside the boundary faces");
        long long onSides = 0;
        for (int a = 0; a < 3; ++a)
            for (int s = 0; s < 2; ++s) {
                const double plane = s ? m.hi[a] : m.lo[a];
                const double tol = 1e-9 * std::max(1.0, std::fabs(plane)) + 1e-12 * spacing[a];
                long long on = 0;
                for (long long f = p.startFace; f < p.startFace + p.nFaces; ++f) on += std::fabs(fc[f][a] - plane) <= tol;
                if (on == 0) continue;
                const int side = 2 * a + s;
                if (!m.sidePatch[side].empty() && m.sidePatch[side] != p.name)
                    throw FatalError(d + ": side " + kSide[side] + " holds patches " + m.sidePatch[side] + " and " +
                                     p.name);
                m.sidePatch[side] = p.name;
                m.sideType[side] = p.type;
                const long long need = nc / m.n[a];
                if (on != need)
                    throw FatalError(d + ": patch " + p.name + " covers " + std::to_string(on) + " of the " +
                                     std::to_string(need) + " faces of side " + kSide[side]);
                onSides += on;
            }
        if (onSides == 0) throw FatalError(d + ": patch " + p.name + " lies on no side of the block (an internal body?)");
        if (onSides != p.nFaces)
            throw FatalError(d + ": patch " + p.name + " has " + std::to_string(p.nFaces - onSides) +
                             " faces inside the block");
    }
    std::string missing;
    for (int s = 0; s < 6; ++s)
        if (m.sidePatch[s].empty()) missing += std::string(missing.empty() ? "" : " ") + kSide[s];
    if (!missing.empty()) throw FatalError(d + ": sides " + missing + " have no patch");
    return m;
}

}  // namespace hs
