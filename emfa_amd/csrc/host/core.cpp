#include "core.h"

#include <sys/stat.h>

#include <cctype>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <map>
#include <array>

#include "emfa_hip.h"

namespace component {

// ------------------------------------------------------------------ config
ConfigLoader::ConfigLoader() { log_ << std::left << "Config Information\n"; }

void ConfigLoader::parse(std::istream &config) {
    auto skip = [&]() {
        for (auto ch = config.peek(); ch != EOF && (std::isspace(ch) || ch == '#'); ch = config.peek()) {
            if (ch == '#')
                config.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
            else
                config.get();
        }
    };
    skip();
    std::string key, value;
    while (config >> key >> value) {
        dict_[key] = Value{false, value};
        skip();
    }
}

void ConfigLoader::load(const std::string &filepath) {
    std::ifstream config(filepath);
    if (config.fail()) throw ConfigInvalid("fail opening config file(" + filepath + ')');
    parse(config);
}

void ConfigLoader::loadFromString(const std::string &text) {
    std::istringstream in(text);
    parse(in);
}

ConfigLoader::Value &ConfigLoader::flag(const char *key) { return dict_.at(key); }

void ConfigLoader::report(const std::string &line) {
    std::cout << line;
    log_ << line;
}

std::string ConfigLoader::getSolverType() {
    auto &v = flag("EMSolver:");
    if (!v.has_reported) {
        v.has_reported = true;
        report("  -> EM Solver: " + v.value + "\n");
    }
    return v.value;
}

policy::ResultType ConfigLoader::getResultType() {
    auto &v = flag("result_type:");
    if (!v.has_reported) {
        v.has_reported = true;
        report("  -> Result type: " + v.value + "\n");
    }
    if (v.value == "BistaticRCS") return policy::BiRCS;
    if (v.value == "MonostaticRCS") return policy::MonoRCS;
    if (v.value == "Radiation") return policy::Rad;
    throw ConfigInvalid(
        "type parameter only supports three type values (BistaticRCS, MonostaticRCS and Radiation)");
}

// std::stof semantics of the reference getters (ConfigLoader.cpp:78-91)
static float to_f(const std::string &s) { return std::stof(s); }

Incidence ConfigLoader::getInc() {
    auto &freq = flag("frequency:");
    auto &ff = flag("frequency_from:");
    auto &ft = flag("frequency_to:");
    auto &fd = flag("frequency_delta:");
    auto &fi = flag("frequency_inc:");
    auto &th = flag("theta:");
    auto &ph = flag("phi:");
    auto &po = flag("pole:");
    if (!freq.has_reported) {
        freq.has_reported = th.has_reported = ph.has_reported = po.has_reported = true;
        report("  -> Incidence:\n    => frequency: " + freq.value + " (Hz)\n    => theta: " + th.value +
               " (degree)\n    => phi: " + ph.value + " (degree)\n    => pole: " + po.value + " (degree)\n");
    }
    Incidence inc;
    try {
        inc.freq = to_f(freq.value);
        inc.freq_from = to_f(ff.value);
        inc.freq_to = to_f(ft.value);
        inc.freq_delta = to_f(fd.value);
        inc.freq_inc = to_f(fi.value);
        inc.theta = to_f(th.value);
        inc.phi = to_f(ph.value);
        inc.pole = to_f(po.value);
    } catch (...) {
        throw ConfigInvalid("fail parsing incidence information");
    }
    return inc;
}

Scattering ConfigLoader::getSca() {
    auto &tf = flag("theta_from:");
    auto &tt = flag("theta_to:");
    auto &td = flag("theta_delta:");
    auto &pf = flag("phi_from:");
    auto &pt = flag("phi_to:");
    auto &pd = flag("phi_delta:");
    if (!tf.has_reported) {
        tf.has_reported = tt.has_reported = td.has_reported = true;
        pf.has_reported = pt.has_reported = pd.has_reported = true;
        report("  -> Scattering:\n    => theta from: " + tf.value + " (degree)\n    => theta to: " + tt.value +
               " (degree)\n    => theta delta: " + td.value + " (degree)\n    => phi from: " + pf.value +
               " (degree)\n    => phi to: " + pt.value + " (degree)\n    => phi delta: " + pd.value + " (degree)\n");
    }
    Scattering sca;
    try {
        sca.theta_from = to_f(tf.value);
        sca.theta_to = to_f(tt.value);
        sca.theta_delta = to_f(td.value);
        sca.phi_from = to_f(pf.value);
        sca.phi_to = to_f(pt.value);
        sca.phi_delta = to_f(pd.value);
    } catch (...) {
        throw ConfigInvalid("fail parsing scattering information");
    }
    return sca;
}

Radiation ConfigLoader::getRad() {
    auto &e = flag("E_Plane:");
    auto &h = flag("H_Plane:");
    auto &hh = flag("HH_Plane:");
    auto tf = [](const std::string &s) {
        if (s == "true") return true;
        if (s == "false") return false;
        throw ConfigInvalid("radiation plane switches must be true or false");
    };
    Radiation r;
    r.e_plane = tf(e.value);
    r.h_plane = tf(h.value);
    r.hh_plane = tf(hh.value);
    return r;
}

// "(x,y,z)" as ConfigLoader.cpp:166-199 parses it
static Vec3 parse_point(const std::string &s, const char *what) {
    auto p1 = s.find_first_of('('), p2 = s.find_first_of(','), p3 = s.find_last_of(','), p4 = s.find_last_of(')');
    Vec3 v;
    try {
        v.x = std::stof(s.substr(p1 + 1, p2 - p1 - 1));
        v.y = std::stof(s.substr(p2 + 1, p3 - p2 - 1));
        v.z = std::stof(s.substr(p3 + 1, p4 - p3 - 1));
    } catch (...) {
        throw ConfigInvalid(std::string("fail parsing ") + what);
    }
    return v;
}

Nearfield ConfigLoader::getNearfield() {
    Nearfield nf;
    nf.origin_point = parse_point(dict_.at("Origin_point:").value, "Origin point");
    nf.end_point = parse_point(dict_.at("End_point:").value, "end point");
    try {
        nf.sampling_x = std::stoi(dict_.at("Sampling_x:").value);
        nf.sampling_y = std::stoi(dict_.at("Sampling_y:").value);
        nf.sampling_z = std::stoi(dict_.at("Sampling_z:").value);
    } catch (const std::out_of_range &) {
        throw;
    } catch (...) {
        throw ConfigInvalid("fail parsing sampling number");
    }
    return nf;
}

std::string ConfigLoader::getMeshPath() {
    auto &v = flag("mesh_path:");
    if (!v.has_reported) {
        v.has_reported = true;
        report("  -> Mesh path: " + v.value + "\n");
    }
    return v.value;
}

#define EMFA_SCALAR_GETTER(NAME, KEY, TYPE, CONV, MSG)      \
    TYPE ConfigLoader::NAME() {                             \
        auto &v = flag(KEY);                                \
        if (!v.has_reported) {                              \
            v.has_reported = true;                          \
            report(std::string("  -> ") + KEY + " " + v.value + "\n"); \
        }                                                   \
        try {                                               \
            return (TYPE)CONV(v.value);                     \
        } catch (...) {                                     \
            throw ConfigInvalid(MSG);                       \
        }                                                   \
    }

EMFA_SCALAR_GETTER(getAlpha, "alpha:", double, std::stof, "fail parsing combined coefficient")
EMFA_SCALAR_GETTER(getIterationThreshold, "iteration_threshold:", double, std::stof, "fail parsing iteration threshold")
EMFA_SCALAR_GETTER(getMaxIterationNum, "max_iteration_number:", int, std::stoi, "fail parsing maximum iteration number")
EMFA_SCALAR_GETTER(getMLFMABoxThreshold, "mlfma_box_threshold:", double, std::stof, "fail parsing MLFMA box threshold")
EMFA_SCALAR_GETTER(getThreadNumber, "thread_number:", size_t, std::stoul, "fail parsing thread number")
EMFA_SCALAR_GETTER(getTaskFactor, "task_factor:", size_t, std::stoul, "fail parsing task factor")
EMFA_SCALAR_GETTER(getDx, "Dx:", double, std::stof, "fail parsing Dx")
EMFA_SCALAR_GETTER(getDy, "Dy:", double, std::stof, "fail parsing Dy")
EMFA_SCALAR_GETTER(getArray_x, "Array_x:", int, std::stoi, "fail parsing Array_x")
EMFA_SCALAR_GETTER(getArray_y, "Array_y:", int, std::stoi, "fail parsing Array_y")
#undef EMFA_SCALAR_GETTER

bool ConfigLoader::getPreconditioningSwitch() {
    auto &v = flag("preconditioning:");
    if (v.value == "true") return true;
    if (v.value == "false") return false;
    throw ConfigInvalid("preconditioning must be true or false");
}

// "(re,im)" as ConfigLoader.cpp:891-911
std::complex<double> ConfigLoader::getSigma() {
    auto &v = flag("Sigma:");
    if (!v.has_reported) {
        v.has_reported = true;
        report("  -> Sigma: " + v.value + "\n");
    }
    const std::string &s = v.value;
    auto p1 = s.find_first_of('('), p2 = s.find_first_of(','), p3 = s.find_first_of(')');
    try {
        float r = std::stof(s.substr(p1 + 1, p2 - p1 - 1));
        float i = std::stof(s.substr(p2 + 1, p3 - p2 - 1));
        return {r, i};
    } catch (...) {
        throw ConfigInvalid("fail parsing sigma");
    }
}

std::string ConfigLoader::getString(const std::string &key, const std::string &dflt) const {
    auto it = dict_.find(key);
    return it == dict_.end() ? dflt : it->second.value;
}
double ConfigLoader::getDouble(const std::string &key, double dflt) const {
    auto it = dict_.find(key);
    if (it == dict_.end()) return dflt;
    try {
        return std::stod(it->second.value);
    } catch (...) {
        throw ConfigInvalid("fail parsing " + key);
    }
}
int ConfigLoader::getInt(const std::string &key, int dflt) const {
    auto it = dict_.find(key);
    if (it == dict_.end()) return dflt;
    try {
        return std::stoi(it->second.value);
    } catch (...) {
        throw ConfigInvalid("fail parsing " + key);
    }
}

// ------------------------------------------------------------------ mesh
static std::string suffix_of(const std::string &f) { return f.substr(f.find_last_of('.') + 1); }

void Mesh::loadMeshFile(const std::string &filename) {
    std::string suf = suffix_of(filename);
    if (suf == "nas") {
        std::ifstream in(filename);
        if (in.fail()) throw FileError("fail loading nas mesh file: " + filename);
        loadNasStream(in);
        return;
    }
    if (suf == "rpt" || suf == "stl") {
        std::ifstream in(filename);
        if (in.fail()) throw FileError("fail loading " + suf + " mesh file: " + filename);
        if (suf == "rpt")
            loadRptStream(in);
        else
            loadStlStream(in);
        return;
    }
    throw FileError("cannot identify mesh file type: " + suf + " (1.rpt, 2.stl(text), 3.nas)");
}

// A coordinate as the reference keeps it after its temp.dat round trip:
// parsed as float, written with the default stream precision (6 significant
// digits), read back as float.
static double round6(float v) {
    std::ostringstream o;
    o << v;
    float r = 0;
    std::istringstream(o.str()) >> r;
    return (double)r;
}

// .rpt of EMComponent/Mesh.cpp:178-218 + :673-724: two header lines, then
// "id n1 n2 n3" per triangle (node positions from 1) until a record does
// not parse, that line and the next skipped, then "id x y z i j" per node in
// position order; anything after the nodes makes the file invalid.
void Mesh::loadRptStream(std::istream &in) {
    const auto maxs = std::numeric_limits<std::streamsize>::max();
    in.ignore(maxs, '\n');
    in.ignore(maxs, '\n');
    std::vector<int> tri;
    int id = 0, p1 = 0, p2 = 0, p3 = 0;
    while (in >> id >> p1 >> p2 >> p3) {
        tri.push_back(p1 - 1);
        tri.push_back(p2 - 1);
        tri.push_back(p3 - 1);
    }
    in.clear();
    in.ignore(maxs, '\n');
    in.ignore(maxs, '\n');
    std::vector<double> xyz;
    float c1 = 0, c2 = 0, c3 = 0;
    while (in >> id >> c1 >> c2 >> c3 >> p1 >> p1) {
        xyz.push_back(round6(c1));
        xyz.push_back(round6(c2));
        xyz.push_back(round6(c3));
    }
    if (!in.eof()) throw FileError("fail loading rpt mesh file: unexpected record after the nodes");
    const int nv = (int)(xyz.size() / 3);
    for (int v : tri)
        if (v < 0 || v >= nv) throw FileError("fail loading rpt mesh file: vertex id out of range");
    setGeometry(std::move(xyz), std::move(tri));
}

// ASCII .stl of EMComponent/Mesh.cpp:220-259 + :781-829: after the first
// line, per facet the three numbers after the next 'l' (the facet normal)
// and the three after each of the next three 'x' (the vertices), until a
// facet's normal does not parse.  The reference keeps each triangle's three
// corners and the file's normal without vertex numbers; here corners with
// identical (rounded) coordinates become one vertex, numbered in order of
// first appearance, so that RWG functions can be built across facets.  The
// normal is recomputed from the corner order, as for the other formats.
void Mesh::loadStlStream(std::istream &in) {
    const auto maxs = std::numeric_limits<std::streamsize>::max();
    in.ignore(maxs, '\n');
    std::vector<double> xyz;
    std::vector<int> tri;
    std::map<std::array<double, 3>, int> ids;
    float c1 = 0, c2 = 0, c3 = 0;
    while (true) {
        in.ignore(maxs, 'l');
        in >> c1 >> c2 >> c3;  // facet normal
        if (in.fail()) break;
        for (int v = 0; v < 3; ++v) {
            in.ignore(maxs, 'x');
            in >> c1 >> c2 >> c3;
            std::array<double, 3> p{round6(c1), round6(c2), round6(c3)};
            auto it = ids.find(p);
            if (it == ids.end()) {
                it = ids.emplace(p, (int)(xyz.size() / 3)).first;
                xyz.insert(xyz.end(), p.begin(), p.end());
            }
            tri.push_back(it->second);
        }
        in.ignore(maxs, 'l');
    }
    if (tri.empty()) throw FileError("fail loading stl mesh file: no facets");
    setGeometry(std::move(xyz), std::move(tri));
}

// NASTRAN dialect of EMComponent/Mesh.cpp:261-309: 9 header lines, then
// "GRID*" records (id x y, continuation with z) and "CTRIA3 id group n1 n2 n3",
// ids consecutive from 1, terminated by a line starting with 'E'.  The
// reference round-trips coordinates through a text file written with the
// default stream precision (6 significant digits) as float
// (Mesh.cpp:283,303 and :756); this reader applies the same rounding so the
// geometry is the one the reference solves.
void Mesh::loadNasStream(std::istream &in) {
    const auto maxs = std::numeric_limits<std::streamsize>::max();
    for (int t = 0; t < 9; ++t) in.ignore(maxs, '\n');
    std::vector<double> xyz;
    std::vector<int> tri;
    int nodeNum = 0, triNum = 0;
    while (true) {
        int ch = in.peek();
        if (ch == EOF) throw FileError("nas file ended before the 'E' record");
        if (ch == 'E') break;
        bool progressed = false;
        while (in.peek() == 'G') {
            int id1 = 0, id2 = 0;
            float c1 = 0, c2 = 0, c3 = 0;
            in.ignore(maxs, '*');
            in >> id1 >> c1 >> c2;
            in.ignore(maxs, '*');
            in >> id2 >> c3;
            if (!in || id2 != ++nodeNum) throw FileError("fail loading nas mesh file: bad GRID record");
            xyz.push_back(round6(c1));
            xyz.push_back(round6(c2));
            xyz.push_back(round6(c3));
            in.get();
            progressed = true;
        }
        while (in.peek() == 'C') {
            int id = 0, grp = 0, a = 0, b = 0, c = 0;
            in.ignore(maxs, ' ');
            in >> id >> grp >> a >> b >> c;
            if (!in || id != ++triNum) throw FileError("fail loading nas mesh file: bad CTRIA3 record");
            tri.push_back(a - 1);
            tri.push_back(b - 1);
            tri.push_back(c - 1);
            in.get();
            progressed = true;
        }
        if (!progressed) throw FileError("fail loading nas mesh file: unexpected record");
    }
    for (int v : tri)
        if (v < 0 || v >= nodeNum) throw FileError("fail loading nas mesh file: vertex id out of range");
    setGeometry(std::move(xyz), std::move(tri));
}

void Mesh::setGeometry(std::vector<double> xyz, std::vector<int> tri) {
    xyz_ = std::move(xyz);
    tri_ = std::move(tri);
}

void Mesh::reportInfo(std::ostream &s) const {
    s << "Mesh Information\n"
      << "  -> Nodes: " << getNodeNum() << '\n'
      << "  -> Triangles: " << getTriangleNum() << '\n';
}

void CommonEdge::buildCommonEdge(const Mesh &mesh) {
    const auto &tri = mesh.tri();
    int nt = (int)(tri.size() / 3), n = 0, nb = 0;
    if (emfa_rwg_build(nt, tri.data(), nullptr, 0, &n, &nb) != EMFA_OK) throw FileError(emfa_last_error());
    flat_.assign(6 * (size_t)std::max(n, 1), 0);
    if (emfa_rwg_build(nt, tri.data(), flat_.data(), n, &n, &nb) != EMFA_OK) throw FileError(emfa_last_error());
    flat_.resize(6 * (size_t)n);
    boundary_ = nb;
    list_.resize(n);
    const auto &xyz = mesh.xyz();
    for (int i = 0; i < n; ++i) {
        const int *r = &flat_[6 * (size_t)i];
        double dx = xyz[3 * r[0]] - xyz[3 * r[1]], dy = xyz[3 * r[0] + 1] - xyz[3 * r[1] + 1],
               dz = xyz[3 * r[0] + 2] - xyz[3 * r[1] + 2];
        list_[i] = {r[0], r[1], r[2], r[3], r[4], r[5], std::sqrt(dx * dx + dy * dy + dz * dz)};
    }
}

void CommonEdge::reportInfo(std::ostream &s) const {
    s << "CommonEdge Information\n"
      << "  -> Total RWGs: " << list_.size() << '\n'
      << "  -> Boundary edges: " << boundary_ << '\n';
}

// ------------------------------------------------------------------ results
namespace result {
void prefixRCSData(std::ostream &strm) {
    strm << std::right << std::fixed << std::setprecision(4);
    strm << std::setw(10) << "Theta\t" << std::setw(10) << "RCS(dBm^2)\n";
}

void prefixSCData(std::ostream &strm) {
    strm << std::fixed << std::setprecision(8) << std::right;
    strm << std::setw(10) << "id" << std::setw(18) << "v1.x" << std::setw(18) << "v1.y" << std::setw(18) << "v1.z"
         << std::setw(18) << "v2.x" << std::setw(18) << "v2.y" << std::setw(18) << "v2.z" << std::setw(18) << "v3.x"
         << std::setw(18) << "v3.y" << std::setw(18) << "v3.z" << std::setw(18) << "magn1" << std::setw(18)
         << "magn2" << std::setw(18) << "magn3" << std::setw(18) << "magn center" << '\n';
}

void writeBistatic(const std::string &dir, const Scattering &sca,
                   const std::function<void(int, const double *, double *)> &rcs) {
    const float d2r = (float)(3.14159265358979323846 / 180.0);
    std::vector<double> dirs;
    std::vector<std::pair<std::string, std::vector<float>>> files;  // name, angle column
    int num = static_cast<int>((sca.phi_to - sca.phi_from) / sca.phi_delta + 1);
    for (int i = 0; i < num; ++i) {
        float phiDegree = sca.phi_from + i * sca.phi_delta;
        float phi = phiDegree * d2r;
        std::string str = std::to_string(phiDegree);
        files.push_back({"/bistatic_rcs_phi_" + str.substr(0, str.find('.')) + ".dat", {}});
        for (float thetaDegree = sca.theta_from; thetaDegree <= sca.theta_to; thetaDegree += sca.theta_delta) {
            float theta = thetaDegree * d2r;
            dirs.insert(dirs.end(), {std::sin(theta) * std::cos(phi), std::sin(theta) * std::sin(phi), std::cos(theta)});
            files.back().second.push_back(thetaDegree);
        }
    }
    files.push_back({"/bistatic_rcs_theta_90.dat", {}});
    for (float phiDegree = 0; phiDegree < 360; phiDegree += 1) {
        float phi = phiDegree * d2r;
        dirs.insert(dirs.end(), {std::cos(phi), std::sin(phi), 0.0});
        files.back().second.push_back(phiDegree);
    }
    const int nd = (int)(dirs.size() / 3);
    std::vector<double> db(nd);
    rcs(nd, dirs.data(), db.data());
    size_t o = 0;
    for (auto &f : files) {
        std::ofstream out(dir + f.first);
        if (out.fail()) throw FileError("cannot write " + dir + f.first);
        prefixRCSData(out);
        for (float ang : f.second) out << std::setw(10) << ang << '\t' << std::setw(10) << (float)db[o++] << '\n';
    }
}

void writeCurrent(const std::string &dir, const std::vector<CurrentData> &data) {
    std::ofstream out(dir + "/current.sc");
    if (out.fail()) throw FileError("cannot write " + dir + "/current.sc");
    out << "triangles: " << data.size() << '\n';
    prefixSCData(out);
    for (size_t t = 0; t < data.size(); ++t) {
        const auto &d = data[t];
        out << std::setw(10) << t + 1 << std::setw(18) << (float)d.v1.x << std::setw(18) << (float)d.v1.y
            << std::setw(18) << (float)d.v1.z << std::setw(18) << (float)d.v2.x << std::setw(18) << (float)d.v2.y
            << std::setw(18) << (float)d.v2.z << std::setw(18) << (float)d.v3.x << std::setw(18) << (float)d.v3.y
            << std::setw(18) << (float)d.v3.z << std::setw(18) << (float)d.magn1 << std::setw(18) << (float)d.magn2
            << std::setw(18) << (float)d.magn3 << std::setw(18) << (float)d.magnc << '\n';
    }
}
}  // namespace result

}  // namespace component

namespace tool {

static std::string formatTime(long long ms) {
    long long s = ms / 1000, m = s / 60, h = m / 60;
    char buf[64];
    std::snprintf(buf, sizeof(buf), "(%lldh %lldm %llds)", h, m % 60, s % 60);
    return buf;
}

long long elapsedTime(Clock &c, std::ostream &strm, const std::string &description) {
    auto now = std::chrono::steady_clock::now();
    long long period = std::chrono::duration_cast<std::chrono::milliseconds>(now - c.prev).count();
    auto old = strm.flags();
    strm << std::left << std::setw(30) << description + ':' << std::right << std::setw(12) << period << " ms"
         << std::setw(15) << formatTime(period) << '\n';
    strm.flags(old);
    c.prev = now;
    return period;
}

long long totalTime(Clock &c, std::ostream &strm) {
    auto now = std::chrono::steady_clock::now();
    long long period = std::chrono::duration_cast<std::chrono::milliseconds>(now - c.start).count();
    auto old = strm.flags();
    strm << std::left << std::setw(30) << "Total time:" << std::right << std::setw(12) << period << " ms"
         << std::setw(15) << formatTime(period) << '\n';
    strm.flags(old);
    return period;
}

std::string creatFolder(const std::string &dir, const std::string &foldername) {
    std::string directory = dir + '/' + foldername;
    ::mkdir(directory.c_str(), 0755);
    return directory;
}

}  // namespace tool
