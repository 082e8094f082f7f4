// OpenFOAM ASCII reader and writer for dsmcFoam+'s lagrangian cloud files.
#include "lagrangianIO.H"
#include "foamText.H"

#include <sys/stat.h>

#include <algorithm>
#include <cctype>
#include <charconv>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace hs {

namespace {

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

std::string cloudDir(const std::string& timeDir) { return timeDir + "/lagrangian/dsmc"; }

void header(std::ostream& o, const char* cls, const std::string& loc, const char* obj) {
    o << "/*--------------------------------*- C++ -*----------------------------------*\\\n"
         "  written by the MI355X DSMC engine (dsmcFoamPlus)\n"
         "\\*---------------------------------------------------------------------------*/\n"
         "FoamFile\n{\n    version     2.0;\n    format      ascii;\n    class       "
      << cls << ";\n    location    \"" << loc << "\";\n    object      " << obj << ";\n}\n"
      << "// * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * //\n\n";
}

const char* kFooter = "\n// ************************************************************************* //\n";

// shortest text that reads back as the same double
struct Real {
    char b[32];
    int n;
    explicit Real(double x) {
        auto r = std::to_chars(b, b + sizeof(b), x);
        n = (int)(r.ptr - b);
    }
};
std::ostream& operator<<(std::ostream& o, const Real& r) { return o.write(r.b, r.n); }

class Out {
public:
    Out(const std::string& dir, const std::string& loc, const char* cls, const char* obj)
        : path_(dir + "/" + obj), f_(path_) {
        if (!f_) throw FatalError("cannot write " + path_);
        header(f_, cls, loc, obj);
    }
    ~Out() noexcept(false) {
        f_ << kFooter;
        f_.close();
        if (!f_ && !std::uncaught_exceptions()) throw FatalError("error writing " + path_);
    }
    std::ofstream& operator()() { return f_; }

private:
    std::string path_;
    std::ofstream f_;
};

}  // namespace

void LagrangianCloud::resize(int64_t n, int m) {
    modes = m;
    px.assign(n, 0.0); py.assign(n, 0.0); pz.assign(n, 0.0);
    ux.assign(n, 0.f); uy.assign(n, 0.f); uz.assign(n, 0.f); erot.assign(n, 0.f);
    cell.assign(n, 0);
    type.assign(n, 0);
    vib.assign((size_t)m * n, 0);
}

bool hasLagrangianCloud(const std::string& timeDir) { return isFile(cloudDir(timeDir) + "/positions"); }

LagrangianCloud readLagrangianCloud(const std::string& timeDir, int modes, int64_t nCells, int nSpecies) {
    const std::string d = cloudDir(timeDir);
    LagrangianCloud c;
    int64_t n = 0;
    {
        FoamText t(d + "/positions");
        n = t.label();
        if (n < 0) t.fail("negative list size");
        if (n > 2000000000) t.fail(std::to_string(n) + " parcels: more than the engine holds (2e9)");
        c.resize(n, modes);
        t.expect('(');
        for (int64_t k = 0; k < n; ++k) {
            t.expect('(');
            c.px[k] = t.real();
            c.py[k] = t.real();
            c.pz[k] = t.real();
            if (!t.peek(')')) t.fail("a position has more than three coordinates (barycentric positions are not read)");
            t.expect(')');
            const long long cl = t.label();
            if (cl < 0 || cl >= nCells) t.fail("cell label " + std::to_string(cl) + " outside the mesh");
            c.cell[k] = (int32_t)cl;
            while (t.atLabel()) t.label();   // tetFacei tetPti of older formats
        }
        t.expect(')');
    }
    auto sized = [&](FoamText& t, int64_t m) {
        if (m != n) t.fail(std::to_string(m) + " entries for " + std::to_string(n) + " parcels");
    };
    {
        FoamText t(d + "/U");
        sized(t, readList<Vec3>(t, vec3, [&](long long k, const Vec3& v) {
            if (k < n) { c.ux[k] = (float)v.x; c.uy[k] = (float)v.y; c.uz[k] = (float)v.z; }
        }));
    }
    if (isFile(d + "/ERot")) {
        FoamText t(d + "/ERot");
        sized(t, readList<double>(t, [](FoamText& u) { return u.real(); },
                                  [&](long long k, double v) { if (k < n) c.erot[k] = (float)v; }));
    }
    if (isFile(d + "/typeId")) {
        FoamText t(d + "/typeId");
        sized(t, readList<long long>(t, [](FoamText& u) { return u.label(); }, [&](long long k, long long v) {
            if (v < 0 || v >= nSpecies) t.fail("typeId " + std::to_string(v) + " outside the typeIdList");
            if (k < n) c.type[k] = (uint8_t)v;
        }));
    }
    if (isFile(d + "/vibLevel")) {
        FoamText t(d + "/vibLevel");
        sized(t, readList<std::vector<long long>>(t, labelList, [&](long long k, const std::vector<long long>& v) {
            if ((int)v.size() > modes)
                t.fail("a vibLevel list of " + std::to_string(v.size()) + " for at most " + std::to_string(modes) +
                       " vibrational modes");
            for (size_t m = 0; m < v.size(); ++m) {
                if (v[m] < 0 || v[m] > 65535) t.fail("vibLevel " + std::to_string(v[m]) + " out of range");
                if (k < n) c.vib[m * n + k] = (uint16_t)v[m];
            }
        }));
    }
    return c;
}

void writeLagrangianCloud(const std::string& timeDir, const std::string& timeName, const LagrangianCloud& c) {
    const std::string d = cloudDir(timeDir);
    mkdirs(d);
    const std::string loc = timeName + "/lagrangian/dsmc";
    const int64_t n = c.size();
    {
        Out o(d, loc, "Cloud<passiveParticle>", "positions");
        o() << n << "\n(\n";
        for (int64_t k = 0; k < n; ++k)
            o() << '(' << Real(c.px[k]) << ' ' << Real(c.py[k]) << ' ' << Real(c.pz[k]) << ") " << c.cell[k] << '\n';
        o() << ")\n";
    }
    {
        Out o(d, loc, "vectorField", "U");
        o() << n << "\n(\n";
        for (int64_t k = 0; k < n; ++k)
            o() << '(' << Real(c.ux[k]) << ' ' << Real(c.uy[k]) << ' ' << Real(c.uz[k]) << ")\n";
        o() << ")\n";
    }
    {
        Out o(d, loc, "scalarField", "ERot");
        o() << n << "\n(\n";
        for (int64_t k = 0; k < n; ++k) o() << Real(c.erot.empty() ? 0.0 : c.erot[k]) << '\n';
        o() << ")\n";
    }
    {
        Out o(d, loc, "labelField", "typeId");
        o() << n << "\n(\n";
        for (int64_t k = 0; k < n; ++k) o() << (c.type.empty() ? 0 : (int)c.type[k]) << '\n';
        o() << ")\n";
    }
    {
        Out o(d, loc, "labelFieldField", "vibLevel");
        o() << n << "\n(\n";
        const int m = std::max(c.modes, 1);
        for (int64_t k = 0; k < n; ++k) {
            o() << m << '(';
            for (int q = 0; q < m; ++q) o() << (q ? " " : "") << (c.modes ? (int)c.vib[(size_t)q * n + k] : 0);
            o() << ")\n";
        }
        o() << ")\n";
    }
    // bookkeeping labels of OpenFOAM's particle: -1 (not injected this step),
    // classification 0, processor 0, and the parcel's index as its id
    for (const auto& u : { std::make_pair("newParcel", -1), std::make_pair("classification", 0),
                           std::make_pair("origProcId", 0) }) {
        Out o(d, loc, "labelField", u.first);
        o() << n << '{' << u.second << "}\n";
    }
    {
        Out o(d, loc, "labelField", "origId");
        o() << n << "\n(\n";
        for (int64_t k = 0; k < n; ++k) o() << k << '\n';
        o() << ")\n";
    }
}

std::vector<double> readVolScalarInternalField(const std::string& path, int64_t n) {
    FoamText t(path);
    t.seek("internalField");
    std::vector<double> v((size_t)n, 0.0);
    if (t.next("uniform")) {
        const double x = t.real();
        for (auto& y : v) y = x;
        return v;
    }
    if (!t.next("nonuniform")) t.fail("internalField is neither uniform nor nonuniform");
    if (!t.next("List<scalar>")) t.fail("internalField is not a List<scalar>");
    const int64_t m = readList<double>(t, [](FoamText& u) { return u.real(); },
                                       [&](long long k, double x) { if (k < n) v[k] = x; });
    if (m != n) t.fail(std::to_string(m) + " values for " + std::to_string(n) + " cells");
    return v;
}

}  // namespace hs
