// dsmcCloud.C -- see dsmcCloud.H.
#include "dsmcCloud.H"

#include <sys/stat.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <dirent.h>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <sstream>

namespace hs {

namespace {

const double kPi = 3.14159265358979323846;

std::string join(const std::vector<std::string>& v) {
    std::string s;
    for (const auto& x : v) s += "    " + x + "\n";
    return s;
}

bool isFile(const std::string& p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

void mkdirs(const std::string& p) {
    std::string cur;
    for (size_t i = 0; i < p.size(); ++i) {
        cur += p[i];
        if (p[i] == '/' || i + 1 == p.size()) mkdir(cur.c_str(), 0755);
    }
}

// --- runtime-selection tables: the names the reference registers ----------
int collisionModelByName(const std::string& name) {
    static const std::map<std::string, int> t = {
        { "NoBinaryCollision", HS_COLL_NONE },
        { "VariableHardSphere", HS_COLL_VHS },
        { "VariableSoftSphere", HS_COLL_VSS },
        { "LarsenBorgnakkeVariableHardSphere", HS_COLL_LB_VHS },
        { "LarsenBorgnakkeVariableSoftSphere", HS_COLL_LB_VSS },
    };
    auto it = t.find(name);
    if (it == t.end()) {
        std::vector<std::string> valid;
        for (auto& kv : t) valid.push_back(kv.first);
        throw FatalError("Unknown BinaryCollisionModel type " + name + ", constructor not in hash table\n\n"
                         "    Valid BinaryCollisionModel types are:\n" + join(valid));
    }
    return it->second;
}

int patchModelByName(const std::string& name) {
    static const std::map<std::string, int> t = {
        { "dsmcDiffuseWallPatch", HS_PATCH_DIFFUSE },
        { "dsmcSpecularWallPatch", HS_PATCH_SPECULAR },
        { "dsmcDeletionPatch", HS_PATCH_DELETION },
    };
    auto it = t.find(name);
    if (it == t.end()) {
        std::vector<std::string> valid;
        for (auto& kv : t) valid.push_back(kv.first);
        throw FatalError("dsmcPatchBoundary::New(): unknown dsmcPatchBoundary type " + name + "\n\n"
                         "    Valid dsmcPatchBoundary types are:\n" + join(valid));
    }
    return it->second;
}

// exact area of [x0,x1]x[y0,y1] inside the disk |p| <= R
double diskRectArea(double x0, double x1, double y0, double y1, double R) {
    const double a = std::max(x0, -R), b = std::min(x1, R);
    if (a >= b) return 0.0;
    std::vector<double> pts = { a, b };
    for (double y : { y0, y1 })
        if (std::fabs(y) < R) {
            const double w = std::sqrt(R * R - y * y);
            for (double x : { -w, w })
                if (a < x && x < b) pts.push_back(x);
        }
    std::sort(pts.begin(), pts.end());
    auto G = [R](double x) {
        x = std::min(std::max(x, -R), R);
        const double s = std::sqrt(std::max((R - x) * (R + x), 0.0));
        return 0.5 * (x * s + R * R * std::atan2(x, s));
    };
    double area = 0.0;
    for (size_t i = 0; i + 1 < pts.size(); ++i) {
        const double lo = pts[i], hi = pts[i + 1];
        if (hi <= lo) continue;
        const double xm = 0.5 * (lo + hi);
        const double h = std::sqrt(std::max(R * R - xm * xm, 0.0));
        const bool topH = h < y1, botH = -h > y0;
        const double top = topH ? h : y1, bot = botH ? -h : y0;
        if (top <= bot) continue;
        const double c = (topH ? 0.0 : y1) - (botH ? 0.0 : y0);
        const double k = (topH ? 1.0 : 0.0) + (botH ? 1.0 : 0.0);
        area += c * (hi - lo) + k * (G(hi) - G(lo));
    }
    return std::min(std::max(area, 0.0), (x1 - x0) * (y1 - y0));
}

const char* kMagic = "HSDSMC02";

template <class T>
void wr(std::ofstream& f, const T* p, size_t n) { f.write((const char*)p, n * sizeof(T)); }
template <class T>
void rd(std::ifstream& f, T* p, size_t n, const std::string& path) {
    f.read((char*)p, n * sizeof(T));
    if (!f) throw FatalError("truncated checkpoint " + path);
}

void writeHeader(std::ostream& o, const std::string& cls, const std::string& loc, const std::string& obj) {
    o << "/*--------------------------------*- C++ -*----------------------------------*\\\n"
         "  written by the MI355X DSMC engine (dsmcFoamPlus)\n"
         "\\*---------------------------------------------------------------------------*/\n"
         "FoamFile\n{\n    version     2.0;\n    format      ascii;\n    class       "
      << cls << ";\n    location    \"" << loc << "\";\n    object      " << obj << ";\n}\n\n";
}

// Values of a field on wall patches, by patch name (faces in engine order).
using PatchValues = std::map<std::string, std::vector<double>>;

std::string patchType(const BlockMesh& m, const std::string& p) {
    for (int s = 0; s < 6; ++s)
        if (m.sidePatch[s] == p && (m.sideType[s] == "empty" || m.sideType[s] == "cyclic")) return m.sideType[s];
    return "calculated";
}

void writeScalarField(const std::string& path, const std::string& time, const std::string& name,
                      const std::string& dims, const std::vector<double>& v, const BlockMesh& m, int prec,
                      const PatchValues* bv = nullptr, const char* wallType = "calculated") {
    std::ofstream o(path);
    if (!o) throw FatalError("cannot write " + path);
    writeHeader(o, "volScalarField", time, name);
    o << "dimensions      " << dims << ";\n\ninternalField   nonuniform List<scalar> \n" << v.size() << "\n(\n";
    o << std::setprecision(prec);
    for (double x : v) o << x << "\n";
    o << ")\n;\n\nboundaryField\n{\n";
    for (const auto& p : m.patchNames) {
        std::string type = patchType(m, p);
        if (type == "calculated") type = wallType;
        o << "    " << p << "\n    {\n        type            " << type << ";\n";
        if (type == "calculated") {
            auto it = bv ? bv->find(p) : PatchValues::const_iterator();
            if (bv && it != bv->end()) {
                o << "        value           nonuniform List<scalar> \n" << it->second.size() << "\n(\n";
                for (double x : it->second) o << x << "\n";
                o << ")\n;\n";
            } else {
                o << "        value           uniform 0;\n";
            }
        }
        o << "    }\n";
    }
    o << "}\n";
}

// bv values are (x y z) triples per face
void writeVectorField(const std::string& path, const std::string& time, const std::string& name,
                      const std::string& dims, const std::vector<double>& x, const std::vector<double>& y,
                      const std::vector<double>& z, const BlockMesh& m, int prec, const PatchValues* bv = nullptr) {
    std::ofstream o(path);
    if (!o) throw FatalError("cannot write " + path);
    writeHeader(o, "volVectorField", time, name);
    o << "dimensions      " << dims << ";\n\ninternalField   nonuniform List<vector> \n" << x.size() << "\n(\n";
    o << std::setprecision(prec);
    for (size_t i = 0; i < x.size(); ++i) o << "(" << x[i] << " " << y[i] << " " << z[i] << ")\n";
    o << ")\n;\n\nboundaryField\n{\n";
    for (const auto& p : m.patchNames) {
        const std::string type = patchType(m, p);
        o << "    " << p << "\n    {\n        type            " << type << ";\n";
        if (type == "calculated") {
            auto it = bv ? bv->find(p) : PatchValues::const_iterator();
            if (bv && it != bv->end()) {
                const auto& w = it->second;
                o << "        value           nonuniform List<vector> \n" << w.size() / 3 << "\n(\n";
                for (size_t i = 0; i + 2 < w.size(); i += 3)
                    o << "(" << w[i] << " " << w[i + 1] << " " << w[i + 2] << ")\n";
                o << ")\n;\n";
            } else {
                o << "        value           uniform (0 0 0);\n";
            }
        }
        o << "    }\n";
    }
    o << "}\n";
}

}  // namespace

// ---------------------------------------------------------------------------
// blockMesh: a single axis-aligned hex block with uniform grading
// ---------------------------------------------------------------------------
BlockMesh readBlockMesh(const dictionary& d) {
    BlockMesh m;
    const double scale = d.found("convertToMeters") ? d.scalar("convertToMeters")
                                                    : d.lookupOrDefault("scale", 1.0);
    std::vector<std::array<double, 3>> v;
    for (const auto& it : d.list("vertices")) {
        auto c = numbers(it, d.name() + "/vertices");
        if (c.size() != 3) throw FatalError(d.name() + ": a vertex needs three coordinates");
        v.push_back({ c[0] * scale, c[1] * scale, c[2] * scale });
    }
    const auto& blocks = d.list("blocks");
    std::vector<NodePtr> bl(blocks.begin(), blocks.end());
    if (bl.empty() || bl[0]->text != "hex") throw FatalError(d.name() + ": expected a hex block");
    int nblocks = 0;
    for (const auto& it : bl)
        if (it->kind == Node::Word && it->text == "hex") ++nblocks;
    if (nblocks != 1)
        throw FatalError(d.name() + ": the engine's mesh is one block; found " + std::to_string(nblocks));
    auto lab = numbers(bl[1], "hex vertex labels");
    if (lab.size() != 8) throw FatalError(d.name() + ": hex needs 8 vertex labels");
    size_t k = 2;
    if (bl[k]->kind == Node::Word) ++k;  // optional zone name
    auto nn = numbers(bl[k], "hex cell counts");
    if (nn.size() != 3) throw FatalError(d.name() + ": hex needs (nx ny nz)");
    for (int a = 0; a < 3; ++a) m.n[a] = (int)nn[a];
    if (k + 2 < bl.size() && bl[k + 1]->text == "simpleGrading") {
        auto g = numbers(bl[k + 2], "simpleGrading");
        for (double x : g)
            if (std::fabs(x - 1.0) > 1e-12) throw FatalError(d.name() + ": only uniform grading (1 1 1) is supported");
    }
    std::array<double, 3> c[8];
    for (int i = 0; i < 8; ++i) {
        const int id = (int)lab[i];
        if (id < 0 || id >= (int)v.size()) throw FatalError(d.name() + ": vertex label out of range");
        c[i] = v[id];
    }
    for (int a = 0; a < 3; ++a) { m.lo[a] = c[0][a]; m.hi[a] = c[6][a]; }
    const int want[8][3] = { { 0, 0, 0 }, { 1, 0, 0 }, { 1, 1, 0 }, { 0, 1, 0 },
                             { 0, 0, 1 }, { 1, 0, 1 }, { 1, 1, 1 }, { 0, 1, 1 } };
    for (int i = 0; i < 8; ++i)
        for (int a = 0; a < 3; ++a) {
            const double x = want[i][a] ? m.hi[a] : m.lo[a];
            if (!(m.hi[a] > m.lo[a]) || std::fabs(c[i][a] - x) > 1e-12 * (1.0 + std::fabs(x)))
                throw FatalError(d.name() + ": the hex block must be an axis-aligned box with vertices in "
                                            "the order (x0y0z0 x1y0z0 x1y1z0 x0y1z0 ...)");
        }
    // patches
    if (d.found("boundary")) {
        const auto& b = d.list("boundary");
        for (size_t i = 0; i + 1 < b.size(); i += 2) {
            const std::string name = b[i]->text;
            if (b[i + 1]->kind != Node::Dict) throw FatalError(d.name() + ": boundary entry " + name + " needs a dictionary");
            dictionary pd(b[i + 1], d.name() + "/boundary/" + name);
            const std::string type = pd.word("type");
            m.patchNames.push_back(name);
            for (const auto& f : pd.list("faces")) {
                auto q = numbers(f, "face");
                int side = -1;
                for (int a = 0; a < 3 && side < 0; ++a)
                    for (int hiSide = 0; hiSide < 2 && side < 0; ++hiSide) {
                        const double x = hiSide ? m.hi[a] : m.lo[a];
                        bool all = true;
                        for (double id : q)
                            if (std::fabs(v[(int)id][a] - x) > 1e-12 * (1.0 + std::fabs(x))) all = false;
                        if (all) side = 2 * a + hiSide;
                    }
                if (side < 0) throw FatalError(d.name() + ": a face of patch " + name + " is not a block side");
                if (!m.sidePatch[side].empty() && m.sidePatch[side] != name)
                    throw FatalError(d.name() + ": a block side belongs to two patches (" + m.sidePatch[side] + ", " +
                                     name + ")");
                m.sidePatch[side] = name;
                m.sideType[side] = type;
            }
        }
    }
    bool defaults = false;
    for (int s = 0; s < 6; ++s)
        if (m.sidePatch[s].empty()) {
            m.sidePatch[s] = "defaultFaces";
            m.sideType[s] = "empty";
            defaults = true;
        }
    if (defaults) m.patchNames.push_back("defaultFaces");
    return m;
}

std::vector<double> cellVolumes(const BlockMesh& m, bool cylinder, double cx, double cy, double r) {
    const double dx = (m.hi[0] - m.lo[0]) / m.n[0], dy = (m.hi[1] - m.lo[1]) / m.n[1],
                 dz = (m.hi[2] - m.lo[2]) / m.n[2];
    const double full = dx * dy * dz;
    std::vector<double> vol((size_t)m.n[0] * m.n[1] * m.n[2], full);
    if (!cylinder) return vol;
    for (int j = 0; j < m.n[1]; ++j)
        for (int i = 0; i < m.n[0]; ++i) {
            const double x0 = m.lo[0] + i * dx - cx, x1 = m.lo[0] + (i + 1) * dx - cx;
            const double y0 = m.lo[1] + j * dy - cy, y1 = m.lo[1] + (j + 1) * dy - cy;
            const double nx = (x0 <= 0 && 0 <= x1) ? 0.0 : std::min(std::fabs(x0), std::fabs(x1));
            const double ny = (y0 <= 0 && 0 <= y1) ? 0.0 : std::min(std::fabs(y0), std::fabs(y1));
            if (nx * nx + ny * ny >= r * r) continue;
            const double frac = 1.0 - diskRectArea(x0, x1, y0, y1, r) / (dx * dy);
            const double v = frac < 1e-12 ? 0.0 : frac * full;
            for (int k = 0; k < m.n[2]; ++k) vol[(size_t)i + (size_t)m.n[0] * (j + (size_t)m.n[1] * k)] = v;
        }
    return vol;
}

// ---------------------------------------------------------------------------
// chemReactDict (dsmcReactions.C:76-116): "reactions ( name { ... } ... );",
// each entry selected by reactionModel (dsmcReaction.C:106-135) with its
// <model>Properties sub-dictionaries
// ---------------------------------------------------------------------------
void dsmcCloud::readReactions(const dictionary& cr) {
    if (!cr.found("reactions")) return;
    const auto& L = cr.list("reactions");
    auto typeId = [&](const std::string& w, const std::string& where) {
        for (size_t k = 0; k < typeIdList_.size(); ++k)
            if (typeIdList_[k] == w) return (int)k;
        throw FatalError(where + ": Cannot find type id: " + w);
    };
    for (size_t i = 0; i + 1 < L.size(); i += 2) {
        const std::string name = L[i]->text;
        if (L[i + 1]->kind != Node::Dict) throw FatalError(cr.name() + ": reaction " + name + " needs a dictionary");
        const dictionary rd(L[i + 1], cr.name() + "/reactions/" + name);
        const std::string model = rd.word("reactionModel");
        hs_reaction x{};
        if (model == "dissociationQK") x.model = HS_REACT_DISSOCIATION_QK;
        else if (model == "exchangeQK") x.model = HS_REACT_EXCHANGE_QK;
        else if (model == "dissociationExchangeQK") x.model = HS_REACT_DISSOCIATION_EXCHANGE_QK;
        else
            throw FatalError(rd.name() + ": unknown dsmc reaction model type " + model +
                             "; valid types are: dissociationQK exchangeQK dissociationExchangeQK");
        const auto reac = rd.wordList("reactants");
        if (reac.size() != 2) throw FatalError(rd.name() + ": There should be two reactants");
        for (int r = 0; r < 2; ++r) x.reactants[r] = typeId(reac[r], rd.name());
        x.allow_splitting = rd.switchOrDefault("allowSplitting", true) ? 1 : 0;
        for (int r = 0; r < 2; ++r) x.diss_products[r][0] = x.diss_products[r][1] = -1;
        x.exch_products[0] = x.exch_products[1] = -1;
        if (x.model != HS_REACT_EXCHANGE_QK) {
            const dictionary dp = rd.subDict("dissociationQKProperties");
            const auto& tok = dp.lookup("dissociationProducts");
            if (tok.size() != 1 || tok[0]->kind != Node::List || tok[0]->items.size() != 2)
                throw FatalError(dp.name() + ": There should be two lists of dissociation products "
                                             "(NB: a list can be left empty)");
            for (int r = 0; r < 2; ++r) {
                const NodePtr& l = tok[0]->items[r];
                if (l->kind != Node::List) throw FatalError(dp.name() + ": dissociationProducts needs lists");
                if (l->items.empty()) continue;
                if (l->items.size() != 2)
                    throw FatalError(dp.name() + ": There should be 2 dissociation products for reactant " + reac[r]);
                for (int k = 0; k < 2; ++k) x.diss_products[r][k] = typeId(l->items[k]->text, dp.name());
            }
        }
        if (x.model != HS_REACT_DISSOCIATION_QK) {
            const dictionary ep = rd.subDict("exchangeQKProperties");
            const auto prods = ep.wordList("exchangeProducts");
            if (prods.size() != 2) throw FatalError(ep.name() + ": There should be two exchange products");
            for (int k = 0; k < 2; ++k) x.exch_products[k] = typeId(prods[k], ep.name());
            x.heat_exchange = ep.scalar("heatOfReactionExchange");
            x.a_coeff = ep.scalar("aCoeff");
            x.b_coeff = ep.scalar("bCoeff");
        }
        reactions_.push_back(x);
    }
    cfg_.reactions = reactions_.empty() ? nullptr : reactions_.data();
    cfg_.n_reactions = (int32_t)reactions_.size();
}

// ---------------------------------------------------------------------------
// construction
// ---------------------------------------------------------------------------
dsmcCloud::dsmcCloud(const std::string& caseDir, bool initialise) : case_(caseDir) {
    const dictionary control = readDictionary(case_ + "/system/controlDict");
    cfg_.dt = control.scalar("deltaT");
    startTime_ = control.lookupOrDefault("startTime", 0.0);
    endTime_ = control.scalar("endTime");
    writeInterval_ = control.scalar("writeInterval");
    writeByStep_ = control.lookupOrDefault("writeControl", std::string("timeStep")) == "timeStep";
    nTerminalOutputs_ = control.lookupOrDefault("nTerminalOutputs", 1);
    precision_ = control.lookupOrDefault("writePrecision", 10);

    mesh_ = hasPolyMesh(case_) ? readPolyMesh(case_) : readBlockMesh(readDictionary(case_ + "/system/blockMeshDict"));
    cfg_.nx = mesh_.n[0];
    cfg_.ny = mesh_.n[1];
    cfg_.nz = mesh_.n[2];
    for (int a = 0; a < 3; ++a) {
        cfg_.origin[a] = mesh_.lo[a];
        cfg_.delta[a] = (mesh_.hi[a] - mesh_.lo[a]) / mesh_.n[a];
    }
    readProperties(readDictionary(case_ + "/constant/dsmcProperties"));
    // embedded cylinder (this engine's cut-cell body, axis z)
    double cx = 0, cy = 0, r = 0;
    if (isFile(case_ + "/system/embeddedCylinderDict")) {
        const dictionary cd = readDictionary(case_ + "/system/embeddedCylinderDict");
        auto c = cd.scalarList("centre");
        if (c.size() < 2) throw FatalError(cd.name() + ": centre needs (x y)");
        cx = c[0];
        cy = c[1];
        r = cd.scalar("radius");
        const std::string model = cd.word("boundaryModel");
        cylPatch_ = cd.lookupOrDefault("patchName", std::string("cylinder"));
        mesh_.patchNames.push_back(cylPatch_);
        cfg_.cyl_enabled = 1;
        cfg_.cyl_center[0] = cx;
        cfg_.cyl_center[1] = cy;
        cfg_.cyl_radius = r;
        cfg_.cyl_patch.model = patchModelByName(model);
        if (cfg_.cyl_patch.model == HS_PATCH_DELETION) throw FatalError(cd.name() + ": the body must be a wall");
        const dictionary props = cd.subOrEmptyDict(model + "Properties");
        if (cfg_.cyl_patch.model == HS_PATCH_DIFFUSE) {
            cfg_.cyl_patch.temperature = props.scalar("temperature");
            auto u = props.scalarListOrDefault("velocity", { 0, 0, 0 });
            for (int a = 0; a < 3; ++a) cfg_.cyl_patch.velocity[a] = u[a];
        }
    }
    vol_ = hs::cellVolumes(mesh_, cfg_.cyl_enabled != 0, cx, cy, r);
    cfg_.cell_volume = vol_.data();
    readBoundaries(readDictionary(case_ + "/system/boundariesDict"));
    if (isFile(case_ + "/system/fieldPropertiesDict")) readFields(readDictionary(case_ + "/system/fieldPropertiesDict"));
    cfg_.flags |= HS_FLAG_SAMPLE;
    // boundaryMeasurements at every diffuse or specular wall
    // (dsmcDiffuseWallPatch.C:167, dsmcSpecularWallPatch.C:81)
    bool walls = cfg_.cyl_enabled != 0;
    for (int sd = 0; sd < 6; ++sd)
        walls = walls || cfg_.patch[sd].model == HS_PATCH_DIFFUSE || cfg_.patch[sd].model == HS_PATCH_SPECULAR;
    if (walls) cfg_.flags |= HS_FLAG_WALL;
    if (isFile(case_ + "/system/chemReactDict")) readReactions(readDictionary(case_ + "/system/chemReactDict"));

    // capacity: the state to restart from, or the fill of dsmcInitialiseDict
    double expect = 0.0, cloudTime = 0.0;
    std::string state, cloudDir;
    LagrangianCloud cloud;
    if (initialise) {
        const dictionary ini = readDictionary(case_ + "/system/dsmcInitialiseDict");
        double vtot = 0.0;
        for (double x : vol_) vtot += x;
        for (const auto& it : ini.list("configurations")) {
            if (it->kind != Node::Dict) continue;
            dictionary c(it, ini.name() + "/configuration");
            const dictionary nd = c.subDict("numberDensities");
            for (const auto& k : nd.toc()) expect += nd.scalar(k) * vtot / cfg_.fn;
        }
    } else {
        // the latest time holding an engine checkpoint or a lagrangian cloud
        // (OpenFOAM's files, e.g. a dsmcFoam+ run's); at one time the
        // checkpoint wins, as it also carries remainders and samples
        double ts = 0.0;
        state = latestTimeWith("/lagrangian/dsmc/hystrath.state", ts);
        cloudDir = latestTimeWith("/lagrangian/dsmc/positions", cloudTime);
        if (!cloudDir.empty() && (state.empty() || cloudTime > ts)) {
            state.clear();
            cloud = readLagrangianCloud(cloudDir, vibModes(), (int64_t)vol_.size(), cfg_.n_species);
            expect = (double)cloud.size();
        } else if (!state.empty()) {
            cloudDir.clear();
            std::ifstream f(state + "/lagrangian/dsmc/hystrath.state", std::ios::binary);
            char magic[8];
            int64_t n = 0;
            f.read(magic, 8);
            f.read((char*)&n, 8);
            if (!f || std::memcmp(magic, kMagic, 8) != 0) throw FatalError("not an engine checkpoint: " + state);
            expect = (double)n;
        } else {
            throw FatalError("no initial parcel state under " + case_ + ": run dsmcInitialisePlus first");
        }
    }
    const double def = std::max(2.0 * expect + 65536.0, 100000.0);
    cfg_.capacity = (int64_t)std::min(def, 2.0e9);
    if (int rc = hs_dsmc_create(&cfg_, &eng_)) {
        throw FatalError(std::string("hs_dsmc_create: ") + hs_dsmc_create_error() + " (" + std::to_string(rc) + ")");
    }
    time_ = startTime_;
    if (!initialise && !state.empty()) readCheckpoint(state + "/lagrangian/dsmc/hystrath.state");
    if (!initialise && !cloudDir.empty()) readLagrangian(cloudDir, cloudTime, cloud);
}

dsmcCloud::~dsmcCloud() {
    if (eng_) hs_dsmc_destroy(eng_);
}

void dsmcCloud::check(int rc, const char* what) const {
    if (rc != HS_OK) throw FatalError(std::string(what) + ": " + hs_dsmc_last_error(eng_));
}

void dsmcCloud::readProperties(const dictionary& d) {
    cfg_.fn = d.scalar("nEquivalentParticles");
    if (d.found("seedNumber")) {
        cfg_.seed = (uint64_t)d.label("seedNumber");
    } else {
        // the reference seeds from the clock when seedNumber is absent (dsmcCloud.C:639-646)
        cfg_.seed = (uint64_t)std::chrono::system_clock::now().time_since_epoch().count();
    }
    const std::string sel = d.lookupOrDefault("collisionPartnerSelectionModel", std::string("noTimeCounter"));
    if (sel != "noTimeCounter")
        throw FatalError("collisionPartnerSelection::New(): unknown or unsupported collisionPartnerSelectionModel " +
                         sel + "\n\n    Valid collisionPartnerSelection types are:\n    noTimeCounter\n");
    const std::string model = d.word("BinaryCollisionModel");
    cfg_.collision_model = collisionModelByName(model);
    const dictionary coeffs = d.subOrEmptyDict(model + "Coeffs");
    cfg_.tref = coeffs.lookupOrDefault("Tref", 273.0);
    cfg_.z_rot = coeffs.lookupOrDefault("rotationalRelaxationCollisionNumber", 5.0);
    cfg_.z_vib_fixed = coeffs.lookupOrDefault("vibrationalRelaxationCollisionNumber", 0.0);
    cfg_.z_el = coeffs.lookupOrDefault("electronicRelaxationCollisionNumber", 500.0);
    const std::string izv = coeffs.lookupOrDefault("inverseZvFormulation", std::string(""));
    cfg_.inv_zv_form = izv == "pre-2008" ? 0 : izv == "2008" ? 1 : 2;
    typeIdList_ = d.wordList("typeIdList");
    if (typeIdList_.empty() || typeIdList_.size() > HS_MAX_SPECIES)
        throw FatalError("typeIdList must name 1 to " + std::to_string(HS_MAX_SPECIES) + " species");
    cfg_.n_species = (int)typeIdList_.size();
    const dictionary mp = d.subDict("moleculeProperties");
    for (size_t i = 0; i < typeIdList_.size(); ++i) {
        const dictionary s = mp.subDict(typeIdList_[i]);
        hs_species& sp = cfg_.species[i];
        sp.mass = s.scalar("mass");
        sp.diameter = s.scalar("diameter");
        sp.omega = s.scalar("omega");
        sp.alpha = s.lookupOrDefault("alpha", 1.0);
        sp.rot_dof = s.lookupOrDefault("rotationalDegreesOfFreedom", 0.0);
        sp.n_vib_modes = (int)s.lookupOrDefault("nVibrationalModes", 0.0);
        sp.charge = s.lookupOrDefault("charge", 0);
        sp.theta_d = s.lookupOrDefault("dissociationTemperature", 0.0);
        const auto th = s.scalarListOrDefault("characteristicVibrationalTemperature", {});
        const auto zr = s.scalarListOrDefault("Zref", {});
        const auto tz = s.scalarListOrDefault("referenceTempForZref", {});
        // dsmcParcelI.H:89-110
        if ((int)th.size() != sp.n_vib_modes)
            throw FatalError("Number of characteristic vibrational temperatures is " + std::to_string(th.size()) +
                             ", instead of " + std::to_string(sp.n_vib_modes));
        if ((int)zr.size() != sp.n_vib_modes)
            throw FatalError("Number of reference vibrational relaxation numbers is" + std::to_string(zr.size()) +
                             ", instead of " + std::to_string(sp.n_vib_modes));
        if ((int)tz.size() != sp.n_vib_modes)
            throw FatalError("Number of reference temperature for vibrational relaxation is" +
                             std::to_string(tz.size()) + ", instead of " + std::to_string(sp.n_vib_modes));
        if (sp.n_vib_modes > HS_MAX_VIB_MODES) throw FatalError("at most 3 vibrational modes are supported");
        for (int m = 0; m < sp.n_vib_modes; ++m) {
            sp.theta_v[m] = th[m];
            sp.zref[m] = zr[m];
            sp.tref_zv[m] = tz[m];
        }
        if (s.lookupOrDefault("nElectronicLevels", 1) != 1)
            throw FatalError("species " + typeIdList_[i] + ": more than one electronic level is not supported");
    }
}

void dsmcCloud::readBoundaries(const dictionary& d) {
    // patch types from the mesh
    for (int s = 0; s < 6; ++s) {
        const std::string& t = mesh_.sideType[s];
        hs_patch& p = cfg_.patch[s];
        p = hs_patch{};
        if (t == "cyclic") p.model = HS_PATCH_CYCLIC;
        else if (t == "empty") p.model = HS_PATCH_EMPTY;
        else if (t == "symmetryPlane" || t == "symmetry") p.model = HS_PATCH_SPECULAR;
        else p.model = -1;  // needs a dsmcPatchBoundaries model
    }
    auto sidesOf = [&](const std::string& name) {
        std::vector<int> s;
        for (int k = 0; k < 6; ++k)
            if (mesh_.sidePatch[k] == name) s.push_back(k);
        if (s.empty()) throw FatalError(d.name() + ": patch " + name + " is not a side of the mesh");
        return s;
    };
    if (d.found("dsmcPatchBoundaries")) {
        for (const auto& it : d.list("dsmcPatchBoundaries")) {
            if (it->kind != Node::Dict) continue;
            dictionary b(it, d.name() + "/dsmcPatchBoundaries/boundary");
            const std::string name = b.subDict("patchBoundaryProperties").word("patchName");
            const std::string model = b.word("boundaryModel");
            const int pm = patchModelByName(model);
            const dictionary props = b.subOrEmptyDict(model + "Properties");
            for (int s : sidesOf(name)) {
                hs_patch& p = cfg_.patch[s];
                if (p.model == HS_PATCH_CYCLIC || p.model == HS_PATCH_EMPTY)
                    throw FatalError(d.name() + ": patch " + name + " is " + mesh_.sideType[s] +
                                     " and takes no patch boundary model");
                p.model = pm;
                if (pm == HS_PATCH_DIFFUSE) {
                    p.temperature = props.scalar("temperature");
                    auto u = props.scalarListOrDefault("velocity", { 0, 0, 0 });
                    for (int a = 0; a < 3; ++a) p.velocity[a] = u[a];
                }
            }
        }
    }
    if (d.found("dsmcGeneralBoundaries")) {
        bool have = false;
        for (const auto& it : d.list("dsmcGeneralBoundaries")) {
            if (it->kind != Node::Dict) continue;
            dictionary b(it, d.name() + "/dsmcGeneralBoundaries/boundary");
            const std::string name = b.subDict("generalBoundaryProperties").word("patchName");
            const std::string model = b.word("boundaryModel");
            if (model != "dsmcFreeStreamInflowPatch")
                throw FatalError("dsmcGeneralBoundary::New(): unknown or unsupported dsmcGeneralBoundary type " + model +
                                 "\n\n    Valid dsmcGeneralBoundary types are:\n    dsmcFreeStreamInflowPatch\n");
            const dictionary props = b.subDict(model + "Properties");
            const double T = props.scalar("translationalTemperature");
            const double Trot = props.lookupOrDefault("rotationalTemperature", T);
            const double Tvib = props.lookupOrDefault("vibrationalTemperature", T);
            auto u = props.scalarList("velocity");
            const dictionary nd = props.subDict("numberDensities");
            double n[HS_MAX_SPECIES] = { 0 };
            for (const auto& sp : props.wordList("typeIds")) {
                auto pos = std::find(typeIdList_.begin(), typeIdList_.end(), sp);
                if (pos == typeIdList_.end()) throw FatalError(props.name() + ": typeId " + sp + " not in typeIdList");
                n[pos - typeIdList_.begin()] = nd.scalar(sp);
            }
            if (have) {
                bool same = cfg_.inflow_t == T && cfg_.inflow_trot == Trot && cfg_.inflow_tvib == Tvib;
                for (int a = 0; a < 3; ++a) same = same && cfg_.inflow_u[a] == u[a];
                for (int s = 0; s < HS_MAX_SPECIES; ++s) same = same && cfg_.inflow_n[s] == n[s];
                if (!same) throw FatalError(d.name() + ": all freestream inflow patches must share one state");
            }
            have = true;
            cfg_.inflow_t = T;
            cfg_.inflow_trot = Trot;
            cfg_.inflow_tvib = Tvib;
            for (int a = 0; a < 3; ++a) cfg_.inflow_u[a] = u[a];
            for (int s = 0; s < HS_MAX_SPECIES; ++s) cfg_.inflow_n[s] = n[s];
            for (int s : sidesOf(name)) cfg_.patch[s].inflow = 1;
        }
    }
    for (int s = 0; s < 6; ++s)
        if (cfg_.patch[s].model < 0)
            throw FatalError("dsmcBoundaries: patch " + mesh_.sidePatch[s] + " (type " + mesh_.sideType[s] +
                             ") has no dsmcPatchBoundaries model");
}

void dsmcCloud::readFields(const dictionary& d) {
    for (const auto& it : d.list("dsmcFields")) {
        if (it->kind != Node::Dict) continue;
        dictionary f(it, d.name() + "/dsmcFields/field");
        const std::string model = f.word("fieldModel");
        if (model != "dsmcVolFields")
            throw FatalError("dsmcField::New(): unknown or unsupported dsmcField type " + model +
                             "\n\n    Valid dsmcField types are:\n    dsmcVolFields\n");
        FieldModel fm;
        const dictionary tp = f.subOrEmptyDict("timeProperties");
        fm.resetAtOutput = tp.switchOrDefault("resetAtOutput", true);
        fm.resetAtOutputUntilTime = tp.lookupOrDefault("resetAtOutputUntilTime", 1e300);
        const dictionary p = f.subDict("dsmcVolFieldsProperties");
        fm.fieldName = p.word("fieldName");
        for (const auto& sp : p.wordList("typeIds")) {
            auto pos = std::find(typeIdList_.begin(), typeIdList_.end(), sp);
            if (pos == typeIdList_.end()) throw FatalError(p.name() + ": typeId " + sp + " not in typeIdList");
            fm.typeIds.push_back((int)(pos - typeIdList_.begin()));
        }
        fm.measureMeanFreePath = p.switchOrDefault("measureMeanFreePath", true);
        fm.measureHeatFluxShearStress = p.switchOrDefault("measureHeatFluxShearStress", false);
        if (fm.measureHeatFluxShearStress) cfg_.flags |= HS_FLAG_HEATFLUX;
        if (fm.measureMeanFreePath) cfg_.flags |= HS_FLAG_COLL_SEP;
        fields_.push_back(fm);
    }
}

// ---------------------------------------------------------------------------
// dsmcInitialise+ (dsmcMeshFill)
// ---------------------------------------------------------------------------
void dsmcCloud::initialise() {
    const dictionary ini = readDictionary(case_ + "/system/dsmcInitialiseDict");
    uint32_t id = 0;
    for (const auto& it : ini.list("configurations")) {
        if (it->kind != Node::Dict) continue;
        dictionary c(it, ini.name() + "/configuration");
        const std::string type = c.lookupOrDefault("type", std::string("dsmcMeshFill"));
        if (type != "dsmcMeshFill")
            throw FatalError("dsmcConfiguration::New(): unknown or unsupported configuration type " + type +
                             "\n\n    Valid types are:\n    dsmcMeshFill\n");
        const dictionary nd = c.subDict("numberDensities");
        double n[HS_MAX_SPECIES] = { 0 };
        for (const auto& sp : nd.toc()) {
            auto pos = std::find(typeIdList_.begin(), typeIdList_.end(), sp);
            if (pos == typeIdList_.end())
                throw FatalError("typeId " + sp + "not defined.");  // dsmcMeshFill.C:130-136
            n[pos - typeIdList_.begin()] = nd.scalar(sp);
        }
        const double T = c.scalar("translationalTemperature");
        const double Trot = c.lookupOrDefault("rotationalTemperature", T);
        const double Tvib = c.lookupOrDefault("vibrationalTemperature", T);
        auto u = c.scalarListOrDefault("velocity", { 0, 0, 0 });
        check(hs_dsmc_fill(eng_, n, T, Trot, Tvib, u.data(), 0, id++), "dsmcMeshFill");
    }
    std::cout << "Initialised " << nParticles() << " parcels" << std::endl;
}

// ---------------------------------------------------------------------------
// time loop
// ---------------------------------------------------------------------------
bool dsmcCloud::loop() {
    const double next = startTime_ + (double)(stepIndex_ + 1) * cfg_.dt;
    if (next > endTime_ + 0.5 * cfg_.dt) return false;
    ++stepIndex_;
    time_ = next;
    return true;
}

bool dsmcCloud::writeTime() const {
    if (writeByStep_) {
        const long long wi = std::max(1LL, (long long)std::llround(writeInterval_));
        return stepIndex_ % wi == 0;
    }
    const double eps = 1e-6 * cfg_.dt;
    return std::floor((time_ + eps - startTime_) / writeInterval_) >
           std::floor((time_ - cfg_.dt + eps - startTime_) / writeInterval_);
}

std::string dsmcCloud::timeName() const {
    std::ostringstream o;
    o << std::setprecision(precision_) << time_;
    return o.str();
}

void dsmcCloud::evolve() { check(hs_dsmc_evolve(eng_, 1), "evolve"); }

int64_t dsmcCloud::nParticles() const {
    int64_t n = 0;
    check(hs_dsmc_particle_count(eng_, &n), "particle count");
    return n;
}

void dsmcCloud::info() {
    const int ns = cfg_.n_species;
    const int64_t nslots = hs_dsmc_sample_slots(&cfg_);
    const size_t nc = vol_.size();
    std::vector<double> s((size_t)nslots * nc);
    int64_t nsampled = 0;
    check(hs_dsmc_read_samples(eng_, s.data(), (int64_t)s.size(), &nsampled), "read samples");
    if (prevSums_.size() != s.size() || nsampled < prevSampled_) prevSums_.assign(s.size(), 0.0);
    double N = 0, lke = 0, er = 0, ev = 0;
    for (int sp = 0; sp < ns; ++sp) {
        const size_t b = HS_S_GLOBAL_COUNT + (size_t)sp * HS_SS_COUNT;
        for (size_t c = 0; c < nc; ++c) {
            N += s[(b + HS_SS_N) * nc + c] - prevSums_[(b + HS_SS_N) * nc + c];
            lke += 0.5 * (s[(b + HS_SS_MCC) * nc + c] - prevSums_[(b + HS_SS_MCC) * nc + c]);
            er += s[(b + HS_SS_EROT) * nc + c] - prevSums_[(b + HS_SS_EROT) * nc + c];
            for (int m = 0; m < HS_MAX_VIB_MODES; ++m)
                ev += s[(b + HS_SS_EVIB0 + m) * nc + c] - prevSums_[(b + HS_SS_EVIB0 + m) * nc + c];
        }
    }
    prevSums_ = s;
    prevSampled_ = nsampled;
    int64_t cnt[9] = { 0 };
    check(hs_dsmc_read_counters(eng_, cnt, 9), "read counters");
    const double fn = cfg_.fn;
    std::cout << "    Number of DSMC particles        = " << nParticles() << "\n";
    if (N > 0) {
        std::cout << "    Average linear kinetic energy   = " << lke / N << "\n"
                  << "    Average rotational energy       = " << er / N << "\n"
                  << "    Average vibrational energy      = " << ev / N << "\n"
                  << "    Average electronic energy       = " << 0.0 << "\n"
                  << "    Total energy                    = " << fn * (lke + er + ev) << "\n";
    }
    std::cout << "    Collisions                      = " << cnt[1] << std::endl;
    if (cnt[5] || cnt[6] || cnt[7])
        throw FatalError("device-side bound tripped since the last report: " + std::to_string(cnt[5]) +
                         " parcels hit the tracking limit, " + std::to_string(cnt[6]) +
                         " cells exceeded the NTC candidate or parcel limit, " + std::to_string(cnt[7]) +
                         " inflow parcels dropped at capacity " + std::to_string(cfg_.capacity));
    check(hs_dsmc_reset_counters(eng_), "reset counters");
}

// ---------------------------------------------------------------------------
// fields, checkpoint
// ---------------------------------------------------------------------------
VolFields computeVolFields(const dsmcCloud& cloud, const std::vector<double>& s, long long nSteps,
                           const FieldModel& fm) {
    const hs_config& cfg = cloud.config();
    const auto& V = cloud.cellVolumes();
    const size_t nc = V.size();
    const double kB = HS_KB, fn = cfg.fn, dt = cfg.dt, T = (double)std::max(nSteps, 1LL);
    auto at = [&](int slot, size_t c) { return s[(size_t)slot * nc + c]; };
    auto sl = [&](int sp, int k) { return HS_S_GLOBAL_COUNT + sp * HS_SS_COUNT + k; };
    VolFields f;
    for (auto* v : { &f.rhoN, &f.rhoM, &f.Ttra, &f.Trot, &f.Tvib, &f.Tov, &f.p, &f.Ma, &f.mfp, &f.mcr,
                     &f.measuredCollisionRate, &f.dsmcNMean, &f.Ux, &f.Uy, &f.Uz })
        v->assign(nc, 0.0);
    const double Tref = 273.0;
    for (size_t c = 0; c < nc; ++c) {
        double N = 0, M = 0, mu[3] = { 0, 0, 0 }, mcc = 0, erot = 0, zrot = 0;
        for (int sp : fm.typeIds) {
            const hs_species& p = cfg.species[sp];
            const double n = at(sl(sp, HS_SS_N), c);
            N += n;
            M += p.mass * n;
            for (int a = 0; a < 3; ++a) mu[a] += at(sl(sp, HS_SS_MUX + a), c);
            mcc += at(sl(sp, HS_SS_MCC), c);
            erot += at(sl(sp, HS_SS_EROT), c);
            zrot += p.rot_dof * n;
        }
        if (!(N > 1e-3) || !(V[c] > 0.0)) {
            f.dsmcNMean[c] = 0.001;
            continue;
        }
        const double nCum = fn * N, mCum = fn * M;
        const double rhoN = nCum / (T * V[c]), rhoM = mCum / (T * V[c]);
        double U[3];
        for (int a = 0; a < 3; ++a) U[a] = fn * mu[a] / mCum;
        const double lkeMean = 0.5 * fn * mcc / (V[c] * T);
        const double Ttra = 2.0 / (3.0 * kB * rhoN) * (lkeMean - 0.5 * rhoM * (U[0] * U[0] + U[1] * U[1] + U[2] * U[2]));
        f.rhoN[c] = rhoN;
        f.rhoM[c] = rhoM;
        f.Ux[c] = U[0];
        f.Uy[c] = U[1];
        f.Uz[c] = U[2];
        f.Ttra[c] = Ttra;
        f.p[c] = rhoN * kB * Ttra;
        f.dsmcNMean[c] = N / T;
        const double Trot = zrot > 1e-15 ? 2.0 * erot / (kB * zrot) : 0.0;
        const double zetaRotTot = zrot / N;
        double Tvib = 0, zvib = 0, mol = 0, molarMass = 0, cv = 0, cp = 0;
        for (int sp : fm.typeIds) {
            const hs_species& p = cfg.species[sp];
            const double ns_ = at(sl(sp, HS_SS_N), c);
            double zs = 0, zt = 0;
            for (int m = 0; m < p.n_vib_modes; ++m) {
                const double ev = at(sl(sp, HS_SS_EVIB0 + m), c);
                if (!(ev > 1e-300 && ns_ > 1e-15)) continue;
                const double im = ev / (kB * p.theta_v[m] * ns_);
                if (!(im > 1e-15)) continue;
                const double lf = std::log(1.0 + 1.0 / im);
                const double tv = p.theta_v[m] / lf, zm = 2.0 * im * lf;
                zs += zm;
                zt = zm * tv;  // dsmcVolFields.C:1462-1463 keeps the last mode's product
            }
            if (zs > 1e-15) {
                mol += fn * ns_;
                Tvib += fn * ns_ * (zt / zs);
                zvib += fn * ns_ * zs;
            }
            const double X = fn * ns_ / nCum;
            molarMass += X * p.mass;
            cv += X * (3.0 + p.rot_dof);
            cp += X * (5.0 + p.rot_dof);
        }
        if (mol > 1e-15) { Tvib /= mol; zvib /= mol; }
        f.Trot[c] = Trot;
        f.Tvib[c] = Tvib;
        f.Tov[c] = (3.0 * Ttra + zetaRotTot * Trot + zvib * Tvib) / (3.0 + zetaRotTot + zvib);
        if (Ttra > 1e-15) f.Ma[c] = std::sqrt(U[0] * U[0] + U[1] * U[1] + U[2] * U[2]) /
                                   std::sqrt(cp / cv * kB / molarMass * Ttra);
        if (fm.measureMeanFreePath && Ttra > 1.0) {
            double mfp = 0, mcr = 0;
            for (int a : fm.typeIds) {
                const hs_species& pa = cfg.species[a];
                double smfp = 0, smcr = 0;
                for (int b : fm.typeIds) {
                    const hs_species& pb = cfg.species[b];
                    const double nb = at(sl(b, HS_SS_N), c);
                    if (!(nb > 1e-15)) continue;
                    const double dpq = 0.5 * (pa.diameter + pb.diameter), om = 0.5 * (pa.omega + pb.omega);
                    const double mr = pa.mass * pb.mass / (pa.mass + pb.mass);
                    const double nq = fn * nb / (V[c] * T);
                    smfp += kPi * dpq * dpq * nq * std::pow(Tref / Ttra, om - 0.5) * std::sqrt(1.0 + pa.mass / pb.mass);
                    smcr += 2.0 * std::sqrt(kPi) * dpq * dpq * nq * std::pow(Ttra / Tref, 1.0 - om) *
                            std::sqrt(2.0 * kB * Tref / mr);
                }
                if (smfp > 1e-15) smfp = 1.0 / smfp;
                const double w = fn * at(sl(a, HS_SS_N), c) / nCum;
                mfp += smfp * w;
                mcr += smcr * w;
            }
            f.mfp[c] = mfp < 1e-15 ? 1e15 : mfp;
            f.mcr[c] = mcr;
            f.measuredCollisionRate[c] = at(HS_S_NCOLL, c) * fn / (nCum * dt);
        }
    }
    return f;
}

WallFields computeWallFields(const dsmcCloud& cloud, const std::vector<double>& W, int64_t nf,
                             const std::vector<double>& normal, long long nSteps, const FieldModel& fm) {
    const hs_config& cfg = cloud.config();
    const double kB = HS_KB, fn = cfg.fn, T = (double)std::max(nSteps, 1LL);
    const double SMALL = 1e-15, VSMALL = 1e-300;
    auto at = [&](int sp, int k, int64_t f) { return W[((size_t)sp * HS_WS_COUNT + k) * nf + f]; };
    WallFields w;
    for (auto* v : { &w.rhoN, &w.rhoM, &w.Ttra, &w.Trot, &w.Tvib, &w.Tov, &w.p, &w.tau, &w.q }) v->assign(nf, 0.0);
    w.U.assign(3 * nf, 0.0);
    w.fD.assign(3 * nf, 0.0);
    for (int64_t f = 0; f < nf; ++f) {
        double rhoNBF = 0, rhoMBF = 0, ke = 0, mom[3] = { 0, 0, 0 }, erot = 0, zrot = 0, q = 0, fd[3] = { 0, 0, 0 };
        for (int sp : fm.typeIds) {
            const hs_species& c = cfg.species[sp];
            const double n = at(sp, HS_WS_RHON, f);
            rhoNBF += n;
            rhoMBF += c.mass * n;
            ke += at(sp, HS_WS_KE, f);
            for (int a = 0; a < 3; ++a) {
                mom[a] += at(sp, HS_WS_MOMX + a, f);
                fd[a] += at(sp, HS_WS_FDX + a, f);
            }
            erot += at(sp, HS_WS_EROT, f);
            zrot += c.rot_dof * n;
            q += at(sp, HS_WS_Q, f);
        }
        const double rhoN = rhoNBF * fn / T, rhoM = rhoMBF * fn / T, lke = ke * fn / T;
        w.rhoN[f] = rhoN;
        w.rhoM[f] = rhoM;
        double U[3] = { 0, 0, 0 };
        if (rhoM > VSMALL) {
            for (int a = 0; a < 3; ++a) U[a] = mom[a] / rhoMBF;
            w.Ttra[f] = 2.0 / (3.0 * kB * rhoN) * (lke - 0.5 * rhoM * (U[0] * U[0] + U[1] * U[1] + U[2] * U[2]));
        }
        const double zetaRotTot = rhoNBF > SMALL ? zrot / rhoNBF : 0.0;
        w.Trot[f] = zrot > SMALL ? 2.0 * erot / (kB * zrot) : 0.0;
        double Tvib = 0.0, zvib = 0.0, mol = 0.0;
        for (int sp : fm.typeIds) {
            const hs_species& c = cfg.species[sp];
            const double ns = at(sp, HS_WS_RHON, f);
            if (!(ns > SMALL)) continue;
            double zs = 0.0, zt = 0.0;
            for (int m = 0; m < c.n_vib_modes; ++m) {
                const double iMean = at(sp, HS_WS_EVIB0 + m, f) / (kB * c.theta_v[m] * ns);
                if (!(iMean > SMALL)) continue;
                const double lf = std::log(1.0 + 1.0 / iMean);
                const double zm = 2.0 * iMean * lf;
                zs += zm;
                zt += zm * c.theta_v[m] / lf;
            }
            if (zs > SMALL) {
                mol += ns;
                Tvib += ns * zt / zs;
                zvib += ns * zs;
            }
        }
        if (mol > SMALL) {
            Tvib /= mol;
            zvib /= mol;
        }
        w.Tvib[f] = Tvib;
        w.Tov[f] = (3.0 * w.Ttra[f] + zetaRotTot * w.Trot[f] + zvib * Tvib) / (3.0 + zetaRotTot + zvib);
        double pn = 0.0, f2 = 0.0;
        for (int a = 0; a < 3; ++a) {
            w.U[3 * f + a] = U[a];
            w.fD[3 * f + a] = fd[a] / T;
            pn += w.fD[3 * f + a] * normal[3 * f + a];
            f2 += w.fD[3 * f + a] * w.fD[3 * f + a];
        }
        w.p[f] = pn;
        w.tau[f] = std::sqrt(std::max(f2 - pn * pn, 0.0));
        w.q[f] = q / T;
    }
    return w;
}

void dsmcCloud::write() {
    const std::string tn = timeName();
    const std::string dir = case_ + "/" + tn;
    mkdirs(dir + "/lagrangian/dsmc");
    const int64_t nslots = hs_dsmc_sample_slots(&cfg_);
    std::vector<double> s((size_t)nslots * vol_.size());
    int64_t nsampled = 0;
    check(hs_dsmc_read_samples(eng_, s.data(), (int64_t)s.size(), &nsampled), "read samples");
    // wall faces (boundaryMeasurements)
    int64_t nw = 0;
    check(hs_dsmc_wall_face_count(eng_, &nw), "wall face count");
    std::vector<int32_t> wside(nw);
    std::vector<double> wnormal(3 * nw), wsums((size_t)cfg_.n_species * HS_WS_COUNT * nw);
    if (nw) {
        check(hs_dsmc_wall_faces(eng_, wside.data(), nullptr, nullptr, wnormal.data()), "wall faces");
        check(hs_dsmc_read_wall_samples(eng_, wsums.data(), (int64_t)wsums.size(), nullptr), "wall samples");
    }
    auto patchOf = [&](int64_t f) { return wside[f] < 6 ? mesh_.sidePatch[wside[f]] : cylPatch_; };
    auto byPatch = [&](const std::vector<double>& v, int width) {
        PatchValues out;
        for (int64_t f = 0; f < nw; ++f) {
            auto& dst = out[patchOf(f)];
            for (int a = 0; a < width; ++a) dst.push_back(v[(size_t)f * width + a]);
        }
        return out;
    };
    bool reset = false;
    for (const auto& fm : fields_) {
        const VolFields f = computeVolFields(*this, s, nsampled, fm);
        // inverseZvFormulation 2008 reads the first field's overall
        // temperature until the next write (fields().overallT, dsmcCloud.C:1440-1455)
        if (&fm == &fields_.front() && cfg_.inv_zv_form == 1 && cfg_.collision_model >= HS_COLL_LB_VHS)
            check(hs_dsmc_set_cell_temperature(eng_, f.Tov.data()), "cell temperature");
        const WallFields w = computeWallFields(*this, wsums, nw, wnormal, nsampled, fm);
        const std::string n = "_" + fm.fieldName;
        const std::vector<double> zero(vol_.size(), 0.0);
        auto wr1 = [&](const char* name, const char* dims, const std::vector<double>& cellv,
                       const std::vector<double>& wallv) {
            const PatchValues bv = byPatch(wallv, 1);
            writeScalarField(dir + "/" + name + n, tn, name + n, dims, cellv, mesh_, precision_, nw ? &bv : nullptr);
        };
        wr1("rhoN", "[0 -3 0 0 0 0 0]", f.rhoN, w.rhoN);
        wr1("rhoM", "[1 -3 0 0 0 0 0]", f.rhoM, w.rhoM);
        wr1("p", "[1 -1 -2 0 0 0 0]", f.p, w.p);
        wr1("Ttra", "[0 0 0 1 0 0 0]", f.Ttra, w.Ttra);
        wr1("Trot", "[0 0 0 1 0 0 0]", f.Trot, w.Trot);
        wr1("Tvib", "[0 0 0 1 0 0 0]", f.Tvib, w.Tvib);
        wr1("Tov", "[0 0 0 1 0 0 0]", f.Tov, w.Tov);
        writeScalarField(dir + "/Ma" + n, tn, "Ma" + n, "[0 0 0 0 0 0 0]", f.Ma, mesh_, precision_);
        writeScalarField(dir + "/dsmcNMean" + n, tn, "dsmcNMean" + n, "[0 0 0 0 0 0 0]", f.dsmcNMean, mesh_, precision_);
        {
            const PatchValues bu = byPatch(w.U, 3);
            writeVectorField(dir + "/U" + n, tn, "U" + n, "[0 1 -1 0 0 0 0]", f.Ux, f.Uy, f.Uz, mesh_, precision_,
                             nw ? &bu : nullptr);
        }
        if (nw) {
            wr1("wallHeatFlux", "[1 0 -3 0 0 0 0]", zero, w.q);
            wr1("wallShearStress", "[1 -1 -2 0 0 0 0]", zero, w.tau);
            const PatchValues bf = byPatch(w.fD, 3);
            writeVectorField(dir + "/fD" + n, tn, "fD" + n, "[1 -1 -2 0 0 0 0]", zero, zero, zero, mesh_, precision_,
                             &bf);
        }
        if (fm.measureMeanFreePath) {
            writeScalarField(dir + "/mfp" + n, tn, "mfp" + n, "[0 1 0 0 0 0 0]", f.mfp, mesh_, precision_);
            writeScalarField(dir + "/meanCollisionRate" + n, tn, "meanCollisionRate" + n, "[0 0 -1 0 0 0 0]", f.mcr,
                             mesh_, precision_);
            writeScalarField(dir + "/measuredCollisionRate" + n, tn, "measuredCollisionRate" + n, "[0 0 -1 0 0 0 0]",
                             f.measuredCollisionRate, mesh_, precision_);
        }
        if (fm.resetAtOutput && time_ < fm.resetAtOutputUntilTime) reset = true;
    }
    writeLagrangian(dir, tn);
    if (reset) check(hs_dsmc_reset_samples(eng_), "reset samples");
    // after the reset: a restart continues with the sums this run continues with
    writeCheckpoint(dir + "/lagrangian/dsmc/hystrath.state");
}

int dsmcCloud::vibModes() const {
    int modes = 0;
    for (int s = 0; s < cfg_.n_species; ++s) modes = std::max(modes, cfg_.species[s].n_vib_modes);
    return modes;
}

bool dsmcCloud::rotational() const {
    for (int s = 0; s < cfg_.n_species; ++s)
        if (cfg_.species[s].rot_dof > 0) return true;
    return false;
}

LagrangianCloud dsmcCloud::downloadCloud() const {
    const int64_t n = nParticles();
    const int ns = cfg_.n_species, modes = vibModes();
    const bool rot = rotational();
    LagrangianCloud c;
    c.resize(n, modes);
    hs_particles p{ c.px.data(), c.py.data(), c.pz.data(), c.ux.data(), c.uy.data(), c.uz.data(), c.cell.data(),
                    rot ? c.erot.data() : nullptr, modes ? c.vib.data() : nullptr, ns > 1 ? c.type.data() : nullptr };
    int64_t got = 0;
    check(hs_dsmc_download_particles(eng_, &p, n, &got), "download particles");
    return c;
}

void dsmcCloud::writeLagrangian(const std::string& timeDir, const std::string& timeName) const {
    writeLagrangianCloud(timeDir, timeName, downloadCloud());
    std::vector<double> sig(vol_.size()), rem(vol_.size());
    check(hs_dsmc_get_cell_state(eng_, sig.data(), rem.data()), "cell state");
    writeScalarField(timeDir + "/dsmcSigmaTcRMax", timeName, "dsmcSigmaTcRMax", "[0 3 -1 0 0 0 0]", sig, mesh_,
                     precision_, nullptr, "zeroGradient");
}

void dsmcCloud::readLagrangian(const std::string& timeDir, double t, LagrangianCloud& c) {
    const size_t nc = vol_.size();
    const std::string sigPath = timeDir + "/dsmcSigmaTcRMax";
    if (!isFile(sigPath)) throw FatalError("cannot find file " + sigPath + " (sigmaTcRMax is read with the cloud)");
    std::vector<double> sig = readVolScalarInternalField(sigPath, (int64_t)nc);
    // collision-selection remainders start uniform in [0, 1) per cell
    // (dsmcCloud.C:105-116), here from a splitmix64 sequence of seedNumber
    std::vector<double> rem(nc);
    uint64_t x = cfg_.seed ^ 0x6a09e667f3bcc909ull;
    for (size_t k = 0; k < nc; ++k) {
        uint64_t z = (x += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        rem[k] = (double)(z >> 11) * 0x1.0p-53;
    }
    const int ns = cfg_.n_species;
    hs_particles p{ c.px.data(), c.py.data(), c.pz.data(), c.ux.data(), c.uy.data(), c.uz.data(), c.cell.data(),
                    rotational() ? c.erot.data() : nullptr, c.modes ? c.vib.data() : nullptr,
                    ns > 1 ? c.type.data() : nullptr };
    check(hs_dsmc_upload_particles(eng_, &p, c.size()), "upload particles");
    check(hs_dsmc_set_cell_state(eng_, sig.data(), rem.data()), "cell state");
    check(hs_dsmc_set_step(eng_, (uint32_t)std::llround(t / cfg_.dt)), "step");
    time_ = startTime_ = t;
    stepIndex_ = 0;
    std::cout << "Read " << c.size() << " parcels at time " << t << " from " << timeDir << "/lagrangian/dsmc"
              << std::endl;
}

void dsmcCloud::writeCheckpoint(const std::string& path) const {
    const LagrangianCloud c = downloadCloud();
    const int64_t got = c.size();
    const int ns = cfg_.n_species, modes = c.modes;
    const bool rot = rotational();
    const std::vector<double>&px = c.px, &py = c.py, &pz = c.pz;
    const std::vector<float>&ux = c.ux, &uy = c.uy, &uz = c.uz, &er = c.erot;
    const std::vector<int32_t>& cell = c.cell;
    const std::vector<uint16_t>& vib = c.vib;
    const std::vector<uint8_t>& type = c.type;
    const size_t nc = vol_.size();
    std::vector<double> sig(nc), rem(nc);
    check(hs_dsmc_get_cell_state(eng_, sig.data(), rem.data()), "cell state");
    int64_t nf = 0;
    hs_dsmc_inflow_face_count(eng_, &nf);
    std::vector<double> acc((size_t)std::max<int64_t>(nf * ns, 1));
    if (nf) check(hs_dsmc_get_inflow_accumulators(eng_, acc.data()), "inflow accumulators");
    uint32_t step = 0;
    check(hs_dsmc_get_step(eng_, &step), "step");
    const int64_t nslots = hs_dsmc_sample_slots(&cfg_);
    std::vector<double> smp((size_t)nslots * nc);
    int64_t nsampled = 0;
    check(hs_dsmc_read_samples(eng_, smp.data(), (int64_t)smp.size(), &nsampled), "samples");
    int64_t nw = 0;
    check(hs_dsmc_wall_face_count(eng_, &nw), "wall faces");
    std::vector<double> wsum((size_t)ns * HS_WS_COUNT * nw);
    if (nw) check(hs_dsmc_read_wall_samples(eng_, wsum.data(), (int64_t)wsum.size(), nullptr), "wall samples");
    std::ofstream f(path, std::ios::binary);
    if (!f) throw FatalError("cannot write " + path);
    f.write(kMagic, 8);
    const int64_t hdr[8] = { got, (int64_t)nc, ns, modes, rot ? 1 : 0, nf, (int64_t)step, nsampled };
    wr(f, hdr, 8);
    wr(f, &time_, 1);
    wr(f, px.data(), got); wr(f, py.data(), got); wr(f, pz.data(), got);
    wr(f, ux.data(), got); wr(f, uy.data(), got); wr(f, uz.data(), got);
    wr(f, cell.data(), got);
    if (rot) wr(f, er.data(), got);
    if (modes) wr(f, vib.data(), (size_t)modes * got);
    if (ns > 1) wr(f, type.data(), got);
    wr(f, sig.data(), nc); wr(f, rem.data(), nc);
    if (nf) wr(f, acc.data(), (size_t)nf * ns);
    wr(f, smp.data(), smp.size());
    wr(f, &nw, 1);
    if (nw) wr(f, wsum.data(), wsum.size());
}

void dsmcCloud::readCheckpoint(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw FatalError("cannot open " + path);
    char magic[8];
    rd(f, magic, 8, path);
    if (std::memcmp(magic, kMagic, 8) != 0) throw FatalError("not an engine checkpoint: " + path);
    int64_t hdr[8];
    rd(f, hdr, 8, path);
    const int64_t n = hdr[0], nc = hdr[1], ns = hdr[2], modes = hdr[3], rot = hdr[4], nf = hdr[5];
    if (nc != (int64_t)vol_.size() || ns != cfg_.n_species) throw FatalError(path + " does not match the case");
    rd(f, &time_, 1, path);
    std::vector<double> px(n), py(n), pz(n);
    std::vector<float> ux(n), uy(n), uz(n), er(rot ? n : 0);
    std::vector<int32_t> cell(n);
    std::vector<uint16_t> vib((size_t)modes * n);
    std::vector<uint8_t> type(ns > 1 ? n : 0);
    rd(f, px.data(), n, path); rd(f, py.data(), n, path); rd(f, pz.data(), n, path);
    rd(f, ux.data(), n, path); rd(f, uy.data(), n, path); rd(f, uz.data(), n, path);
    rd(f, cell.data(), n, path);
    if (rot) rd(f, er.data(), n, path);
    if (modes) rd(f, vib.data(), (size_t)modes * n, path);
    if (ns > 1) rd(f, type.data(), n, path);
    std::vector<double> sig(nc), rem(nc), acc((size_t)std::max<int64_t>(nf * ns, 1));
    rd(f, sig.data(), nc, path);
    rd(f, rem.data(), nc, path);
    if (nf) rd(f, acc.data(), (size_t)nf * ns, path);
    hs_particles p{ px.data(), py.data(), pz.data(), ux.data(), uy.data(), uz.data(), cell.data(),
                    rot ? er.data() : nullptr, modes ? vib.data() : nullptr, ns > 1 ? type.data() : nullptr };
    check(hs_dsmc_upload_particles(eng_, &p, n), "upload particles");
    check(hs_dsmc_set_cell_state(eng_, sig.data(), rem.data()), "cell state");
    if (nf) check(hs_dsmc_set_inflow_accumulators(eng_, acc.data()), "inflow accumulators");
    check(hs_dsmc_set_step(eng_, (uint32_t)hdr[6]), "step");
    // the cumulative cell and wall sums the run had after its last write
    const int64_t nslots = hs_dsmc_sample_slots(&cfg_);
    std::vector<double> smp((size_t)nslots * nc);
    rd(f, smp.data(), smp.size(), path);
    int64_t nw = 0, nw_case = 0;
    rd(f, &nw, 1, path);
    check(hs_dsmc_wall_face_count(eng_, &nw_case), "wall faces");
    if (nw != nw_case) throw FatalError(path + " does not match the case's wall faces");
    std::vector<double> wsum((size_t)ns * HS_WS_COUNT * nw);
    if (nw) rd(f, wsum.data(), wsum.size(), path);
    check(hs_dsmc_set_samples(eng_, smp.data(), (int64_t)smp.size(), nw ? wsum.data() : nullptr,
                              (int64_t)wsum.size(), hdr[7]),
          "restore samples");
    startTime_ = time_;
    stepIndex_ = 0;
    std::cout << "Read " << n << " parcels at time " << time_ << " from " << path << std::endl;
}

std::string dsmcCloud::latestTimeWith(const std::string& rel, double& t) const {
    DIR* d = opendir(case_.c_str());
    if (!d) throw FatalError("cannot open case directory " + case_);
    double best = -1e300;
    std::string bestDir;
    while (dirent* e = readdir(d)) {
        const std::string name = e->d_name;
        char* end = nullptr;
        const double tn = std::strtod(name.c_str(), &end);
        if (name.empty() || *end != '\0') continue;
        const std::string p = case_ + "/" + name;
        if (isFile(p + rel) && tn > best) {
            best = tn;
            bestDir = p;
        }
    }
    closedir(d);
    t = bestDir.empty() ? 0.0 : best;
    return bestDir;
}

}  // namespace hs
