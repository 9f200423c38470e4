// Host-side components of the drop-in: configuration, meshes, results.
// Names follow the reference's namespaces (component::, policy::) so solver
// code reads like EMFA's.
#pragma once
#include <array>
#include <chrono>
#include <complex>
#include <fstream>
#include <map>
#include <ostream>
#include <sstream>
#include <stdexcept>
#include <functional>
#include <string>
#include <vector>

namespace component {

// EMComponent/ErrorHandle.h:10-20
class ConfigInvalid : public std::invalid_argument {
  public:
    explicit ConfigInvalid(const std::string &w) : std::invalid_argument(w) {}
};
class FileError : public std::runtime_error {
  public:
    explicit FileError(const std::string &w) : std::runtime_error(w) {}
};

struct Vec3 {
    double x = 0, y = 0, z = 0;
};

// EMComponent/fieldStruct.h
struct Incidence {
    double freq = 0, theta = 0, phi = 0, pole = 0;
    double freq_from = 0, freq_to = 0, freq_delta = 0, freq_inc = 0;
};
struct Scattering {
    float theta_from = 0, theta_to = 0, theta_delta = 1;
    float phi_from = 0, phi_to = 0, phi_delta = 1;
};
struct Radiation {
    bool e_plane = false, h_plane = false, hh_plane = false;
};
struct Nearfield {
    Vec3 origin_point, end_point;
    int sampling_x = 1, sampling_y = 1, sampling_z = 1;
};

}  // namespace component

namespace policy {
enum ResultType { BiRCS = 0, MonoRCS, Rad };  // EMComponent/Custom.h:206
}

namespace component {

// Whitespace-tokenised "key: value" map with '#' comments at token
// boundaries, same semantics as EMComponent/ConfigLoader.cpp:23-50 (keys keep
// their colon, later duplicates win, unknown keys are accepted).  Getters
// throw std::out_of_range for a missing key (dict_.at) and ConfigInvalid for
// an unparsable value, and report each group once.
class ConfigLoader {
  public:
    ConfigLoader();
    ConfigLoader(const ConfigLoader &) = delete;
    ConfigLoader &operator=(const ConfigLoader &) = delete;

    void load(const std::string &filepath);
    void loadFromString(const std::string &text);

    std::string getSolverType();
    policy::ResultType getResultType();
    Incidence getInc();
    Scattering getSca();
    Radiation getRad();
    Nearfield getNearfield();
    std::string getMeshPath();
    double getAlpha();
    std::complex<double> getSigma();
    double getIterationThreshold();
    int getMaxIterationNum();
    double getMLFMABoxThreshold();
    bool getPreconditioningSwitch();
    size_t getThreadNumber();
    size_t getTaskFactor();
    double getDx();
    double getDy();
    int getArray_x();
    int getArray_y();

    // keys of this build (optional; defaults when absent)
    std::string getString(const std::string &key, const std::string &dflt) const;
    double getDouble(const std::string &key, double dflt) const;
    int getInt(const std::string &key, int dflt) const;
    bool has(const std::string &key) const { return dict_.count(key) != 0; }
    const std::string &raw(const std::string &key) const { return dict_.at(key).value; }

    std::string getReport() const { return log_.str(); }
    size_t size() const { return dict_.size(); }

  private:
    struct Value {
        bool has_reported = false;
        std::string value;
    };
    void parse(std::istream &in);
    void report(const std::string &line);
    Value &flag(const char *key);
    std::map<std::string, Value> dict_;
    std::ostringstream log_;
};

// Triangle mesh: vertices (metres) and 0-based triangles.
class Mesh {
  public:
    // EMComponent/Mesh.cpp:22-49: chosen by suffix (.nas, .rpt, .stl), parsed in memory
    void loadMeshFile(const std::string &filename);
    void loadNasStream(std::istream &in);
    void loadRptStream(std::istream &in);
    void loadStlStream(std::istream &in);
    void setGeometry(std::vector<double> xyz, std::vector<int> tri);
    size_t getNodeNum() const { return xyz_.size() / 3; }
    size_t getTriangleNum() const { return tri_.size() / 3; }
    const std::vector<double> &xyz() const { return xyz_; }
    const std::vector<int> &tri() const { return tri_; }
    void reportInfo(std::ostream &s) const;
    void clear() {
        xyz_.clear();
        tri_.clear();
    }

  private:
    std::vector<double> xyz_;
    std::vector<int> tri_;
};

// RWG list in the reference's order (built by the library, see emfa_rwg_build).
class CommonEdge {
  public:
    struct SCE {
        int v1, v2, vxp, vxm, tpid, tmid;
        double length;
    };
    void buildCommonEdge(const Mesh &mesh);
    int getCommonEdgeNum() const { return (int)list_.size(); }
    const SCE &getRWGRef(size_t i) const { return list_[i]; }
    const std::vector<int> &flat() const { return flat_; }
    int boundaryEdges() const { return boundary_; }
    void reportInfo(std::ostream &s) const;
    void clear() {
        list_.clear();
        flat_.clear();
    }

  private:
    std::vector<SCE> list_;
    std::vector<int> flat_;
    int boundary_ = 0;
};

struct CurrentData {
    Vec3 v1, v2, v3;
    double magn1 = 0, magn2 = 0, magn3 = 0, magnc = 0;
};

// Writers with the file names and formats of EMComponent/Result.h:48-155,
// 234-255 and Result.cpp:26-68.
namespace result {
void prefixRCSData(std::ostream &s);
void prefixSCData(std::ostream &s);
void writeCurrent(const std::string &dir, const std::vector<CurrentData> &data);
// The bistatic files of EMComponent/Result.h:48-104 (bistatic_rcs_phi_<phi>.dat
// per phi cut, then bistatic_rcs_theta_90.dat); rcs(nd, dirs[nd][3], out_db)
// gives the values of all directions at once, in file order.
void writeBistatic(const std::string &dir, const Scattering &sca,
                   const std::function<void(int, const double *, double *)> &rcs);
}  // namespace result

}  // namespace component

namespace tool {
// EMComponent/tools.cpp:43-76
struct Clock {
    std::chrono::steady_clock::time_point start = std::chrono::steady_clock::now(), prev = start;
};
long long elapsedTime(Clock &c, std::ostream &strm, const std::string &description);
long long totalTime(Clock &c, std::ostream &strm);
std::string creatFolder(const std::string &dir, const std::string &foldername);
}  // namespace tool
