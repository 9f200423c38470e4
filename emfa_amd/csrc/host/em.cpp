#include "em.h"

#include <cmath>
#include <cstdio>
#include <iomanip>
#include <iostream>
#include <mutex>

namespace mom {

static const double kPi = 3.14159265358979323846264338327950288;
static const double kC0 = 299792458.0;

// ------------------------------------------------------------------ EM
EM::EM() {
    logger_ << std::left;
    time_logger_ << std::left;
}

EM::~EM() {
    tool::totalTime(clock_, time_logger_);
    std::string t = "Time Information\n" + time_logger_.str();
    std::cout << t << std::flush;
    if (logger_.is_open()) logger_ << t << std::flush;
}

void EM::timeLog(const std::string &what) { tool::elapsedTime(clock_, time_logger_, "Time of " + what); }

// EMSolver/EM.cpp:24-61
void EM::init(component::ConfigLoader *ploader) {
    rtype_ = ploader->getResultType();
    incidence_ = ploader->getInc();
    const double d2r = kPi / 180.0;
    double pol = incidence_.pole * d2r, theta = incidence_.theta * d2r, phi = incidence_.phi * d2r;
    inc_k_ = {-std::sin(theta) * std::cos(phi), -std::sin(theta) * std::sin(phi), -std::cos(theta)};
    inc_e_ = {-std::cos(pol) * std::cos(theta) * std::cos(phi) - std::sin(pol) * std::sin(phi),
              -std::cos(pol) * std::cos(theta) * std::sin(phi) + std::sin(pol) * std::cos(phi),
              std::cos(pol) * std::sin(theta)};
    inc_h_ = {inc_k_[1] * inc_e_[2] - inc_k_[2] * inc_e_[1], inc_k_[2] * inc_e_[0] - inc_k_[0] * inc_e_[2],
              inc_k_[0] * inc_e_[1] - inc_k_[1] * inc_e_[0]};
    lamda = kC0 / incidence_.freq;
    if (rtype_ == policy::Rad)
        radiation_ = ploader->getRad();
    else
        scattering_ = ploader->getSca();
    nearfield_ = ploader->getNearfield();
    std::string mesh_file = ploader->getMeshPath();
    auto pos_suffix = mesh_file.find_last_of('.');
    if (pos_suffix == std::string::npos) throw component::ConfigInvalid("mesh path parameter cannot match");
    auto pos1 = mesh_file.find_last_of('/');
    auto pos2 = mesh_file.find_last_of('\\');
    if (pos1 == std::string::npos && pos2 == std::string::npos) {
        dir_ = ".";
        folder_name_ = mesh_file.substr(0, pos_suffix);
    } else {
        auto pos_dir = (pos1 != std::string::npos && pos2 != std::string::npos) ? std::max(pos1, pos2)
                                                                                : std::min(pos1, pos2);
        dir_ = mesh_file.substr(0, pos_dir);
        folder_name_ = mesh_file.substr(pos_dir + 1, pos_suffix - pos_dir - 1);
    }
}

// ------------------------------------------------------------------ device
emfa_ctx *device_context(int device) {
    static std::mutex mu;
    static emfa_ctx *ctx = nullptr;
    static int dev = -1;
    std::lock_guard<std::mutex> lock(mu);
    if (ctx && dev == device) return ctx;
    if (ctx) emfa_ctx_destroy(ctx);
    ctx = nullptr;
    emfa_status st = emfa_ctx_create(device, &ctx);
    if (st != EMFA_OK) throw std::runtime_error(std::string("device: ") + emfa_last_error());
    dev = device;
    return ctx;
}

// ------------------------------------------------------------------ DenseMoM
DenseMoM::~DenseMoM() { clear(); }

void DenseMoM::check(emfa_status st, const char *what) const {
    if (st == EMFA_OK) return;
    std::string msg = std::string(what) + ": " + emfa_last_error();
    if (st == EMFA_E_INVALID) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

void DenseMoM::init(component::ConfigLoader *ploader) {
    EM::init(ploader);
    if (rtype_ == policy::Rad)
        throw component::ConfigInvalid("result_type Radiation (delta-gap .rad excitation) is not part of this build");
    dir_ = tool::creatFolder(dir_, folder_name_ + folderTag());
    logger_.open(dir_ + "/sca_runtime.log");
    mesh_.loadMeshFile(ploader->getMeshPath());
    ce_.buildCommonEdge(mesh_);
    unknowns_ = ce_.getCommonEdgeNum();
    if (unknowns_ <= 0) throw component::FileError("mesh has no interior edges (no RWG unknowns)");
    k_ = 2.0 * kPi * incidence_.freq / kC0;
    if (fillKind() == EMFA_EFIE && ploader->has("Sigma:")) sigma_ = ploader->getSigma();
    std::string p = ploader->getString("precision:", "FP64");
    if (p == "FP64")
        prec_ = EMFA_F64;
    else if (p == "FP32")
        prec_ = EMFA_F32;
    else
        throw component::ConfigInvalid("precision must be FP64 or FP32");
    quad_ = ploader->getInt("quadrature_order:", 3);
    if (quad_ != 3 && quad_ != 7) throw component::ConfigInvalid("quadrature_order must be 3 or 7");
    std::string ls = ploader->getString("linear_solver:", defaultSolver());
    if (ls == "LU")
        solver_ = LinearSolver::LU;
    else if (ls == "GMRES")
        solver_ = LinearSolver::GMRES;
    else if (ls == "BICGSTAB")
        solver_ = LinearSolver::BICGSTAB;
    else
        throw component::ConfigInvalid("linear_solver must be LU, GMRES or BICGSTAB");
    gmres_m_ = ploader->getInt("gmres_restart:", 30);
    gmres_tol_ = ploader->getDouble("gmres_tolerance:", 1e-6);
    bicg_tol_ = ploader->getDouble("bicgstab_tolerance:", 0.01);
    bicg_max_ = ploader->getInt("bicgstab_max_iterations:", 500);
    if (!(bicg_tol_ > 0) || bicg_max_ < 1)
        throw component::ConfigInvalid("bicgstab_tolerance and bicgstab_max_iterations must be positive");
    std::string cm = ploader->getString("cfie_self_term:", "reference");
    if (cm == "reference")
        cfie_mode_ = 0;
    else if (cm == "accurate")
        cfie_mode_ = 1;
    else
        throw component::ConfigInvalid("cfie_self_term must be reference or accurate");
    write_matrix_ = ploader->getString("write_matrix:", "false") == "true";
    device_ = ploader->getInt("device:", 0);

    // right-hand sides: one incidence (bistatic) or one per direction (monostatic)
    rhs_angles_.clear();
    rcs_dirs_.clear();
    if (rtype_ == policy::BiRCS) {
        rhs_angles_.push_back({incidence_.theta, incidence_.phi, incidence_.pole});
    } else {
        // angle order of EMComponent/Result.h:107-155
        const auto &sca = scattering_;
        int phi_num = static_cast<int>((sca.phi_to - sca.phi_from) / sca.phi_delta) + 1;
        for (int i = 0; i < phi_num; ++i) {
            float phiDegree = sca.phi_from + i * sca.phi_delta;
            for (float thetaDegree = sca.theta_from; thetaDegree <= sca.theta_to; thetaDegree += sca.theta_delta)
                rhs_angles_.push_back({thetaDegree, phiDegree, incidence_.pole});
        }
        for (float phiDegree = 0; phiDegree < 360; phiDegree += 1) rhs_angles_.push_back({90.0, phiDegree, incidence_.pole});
    }
    logger_ << ploader->getReport();
    reportInfo(logger_);
    timeLog("init");
}

const char *DenseMoM::solverName(LinearSolver s) {
    return s == LinearSolver::LU ? "LU" : s == LinearSolver::GMRES ? "GMRES" : "BICGSTAB";
}

void DenseMoM::reportInfo(std::ostream &strm) const {
    mesh_.reportInfo(strm);
    ce_.reportInfo(strm);
    strm << "Device path\n"
         << "  -> precision: " << (prec_ == EMFA_F64 ? "FP64" : "FP32 + FP64 refinement") << '\n'
         << "  -> quadrature: " << quad_ << "x" << quad_ << '\n'
         << "  -> linear solver: " << solverName(solver_) << '\n'
         << "  -> right-hand sides: " << rhs_angles_.size() << '\n';
}

void DenseMoM::solve() {
    std::cout << "Solve Information\n";
    ctx_ = device_context(device_);
    const auto &xyz = mesh_.xyz();
    const auto &tri = mesh_.tri();
    check(emfa_mesh_upload(ctx_, xyz.data(), (int)(xyz.size() / 3), tri.data(), (int)(tri.size() / 3),
                           ce_.flat().data(), unknowns_, &dmesh_),
          "mesh upload");
    const int n = unknowns_;
    ncols_ = (int)rhs_angles_.size();
    emfa_mat *Z = nullptr, *B = nullptr;
    emfa_lu *lu = nullptr;
    auto release = [&]() {
        if (lu) emfa_lu_destroy(lu);
        if (Z) emfa_mat_destroy(Z);
        if (B) emfa_mat_destroy(B);
        lu = nullptr;
        Z = B = nullptr;
    };
    try {
        std::vector<double> inc(9 * (size_t)ncols_);
        const double d2r = kPi / 180.0;
        for (int j = 0; j < ncols_; ++j) {
            double th = rhs_angles_[j][0] * d2r, ph = rhs_angles_[j][1] * d2r, po = rhs_angles_[j][2] * d2r;
            double kv[3] = {-std::sin(th) * std::cos(ph), -std::sin(th) * std::sin(ph), -std::cos(th)};
            double ev[3] = {-std::cos(po) * std::cos(th) * std::cos(ph) - std::sin(po) * std::sin(ph),
                            -std::cos(po) * std::cos(th) * std::sin(ph) + std::sin(po) * std::cos(ph),
                            std::cos(po) * std::sin(th)};
            double hv[3] = {kv[1] * ev[2] - kv[2] * ev[1], kv[2] * ev[0] - kv[0] * ev[2],
                            kv[0] * ev[1] - kv[1] * ev[0]};
            for (int c = 0; c < 3; ++c) {
                inc[9 * j + c] = kv[c];
                inc[9 * j + 3 + c] = ev[c];
                inc[9 * j + 6 + c] = hv[c];
            }
            rcs_dirs_.push_back({-kv[0], -kv[1], -kv[2]});
        }
        check(emfa_mat_create(ctx_, n, ncols_, EMFA_F64, &B), "excitation");
        check(emfa_rhs_fill(ctx_, dmesh_, rhsKind(), k_, alpha(), ncols_, inc.data(), B), "fillV");
        timeLog("fillV");

        check(emfa_mat_create(ctx_, n, ncols_, EMFA_F64, &X_), "solution");
        const bool matrix_free = solveMatrixFree(B);
        if (!matrix_free) {
            check(emfa_mat_create(ctx_, n, n, prec_, &Z), "impedance matrix");
            emfa_fill_params fp{fillKind(), quad_, k_, alpha(), sigma_.real(), sigma_.imag(), cfie_mode_};
            check(emfa_fill(ctx_, dmesh_, &fp, Z), "fill");
            check(emfa_ctx_sync(ctx_), "fill");
            timeLog("fillZ");
        }
        if (matrix_free) {
        } else if (solver_ == LinearSolver::BICGSTAB) {
            // EMSolver/CFIE.cpp:74-80: iml::BICGSTAB(Z, V, I, 0.01f, 500, iter_num)
            emfa_bicgstab_report rep{};
            emfa_status st = emfa_bicgstab(ctx_, Z, B, X_, bicg_tol_, bicg_max_, 0, &rep);
            if (st == EMFA_E_NOCONV)
                logger_ << "  -> WARNING: Iteration cannot converge\n";
            else
                check(st, "BICGSTAB");
            logger_ << "  -> Iterative number: " << rep.iterations << ", relative residual: " << rep.relres << '\n';
            std::cout << "  -> Iterative number: " << rep.iterations << std::endl;
        } else if (solver_ == LinearSolver::GMRES) {
            emfa_gmres_report rep{};
            emfa_status st = emfa_gmres(ctx_, Z, B, X_, gmres_m_, gmres_tol_, 1000, &rep);
            logger_ << "  -> GMRES iterations: " << rep.iterations << ", restarts: " << rep.restarts
                    << ", relative residual: " << rep.relres << '\n';
            std::cout << "  -> GMRES iterations: " << rep.iterations << '\n';
            if (st == EMFA_E_NOCONV)
                logger_ << "  -> WARNING: Iteration cannot converge\n";
            else
                check(st, "GMRES");
        } else {
            emfa_status st = emfa_lu_factor(ctx_, Z, prec_, &lu);
            if (st == EMFA_E_SINGULAR) throw std::runtime_error("fail solving matrix equation (singular matrix)");
            check(st, "LU");
            if (prec_ == EMFA_F64) {
                check(emfa_mat_copy(X_, B), "solve");
                check(emfa_lu_solve(ctx_, lu, X_), "solve");
            } else {
                emfa_refine_report rep{};
                emfa_status rs = emfa_refine(ctx_, Z, lu, B, X_, 1e-12, 10, &rep);
                logger_ << "  -> refinement steps: " << rep.iterations << ", relative residual " << rep.relres_first
                        << " -> " << rep.relres_final << '\n';
                if (rs != EMFA_E_NOCONV) check(rs, "refinement");
            }
        }
        check(emfa_ctx_sync(ctx_), "solve");
        I_.assign((size_t)n * ncols_, {0, 0});
        check(emfa_mat_download(X_, reinterpret_cast<double *>(I_.data())), "solution download");
        timeLog("solve");
        if (write_matrix_ && Z) {
            std::vector<std::complex<double>> Zh((size_t)n * n), Bh((size_t)n * ncols_);
            check(emfa_mat_download(Z, reinterpret_cast<double *>(Zh.data())), "Z download");
            check(emfa_mat_download(B, reinterpret_cast<double *>(Bh.data())), "V download");
            auto save = [&](const std::string &name, const std::vector<std::complex<double>> &m, int rows, int cols) {
                // arma_ascii as EFIE.cpp:932-950 (complex<float> matrices)
                std::ofstream f(dir_ + "/" + name);
                f << "ARMA_MAT_TXT_FC008\n" << rows << ' ' << cols << '\n';
                for (int r = 0; r < rows; ++r) {
                    for (int c = 0; c < cols; ++c) {
                        std::ostringstream ss;
                        ss << '(' << (float)m[(size_t)r * cols + c].real() << ',' << (float)m[(size_t)r * cols + c].imag()
                           << ')';
                        f.put(' ');
                        f << ss.str();
                    }
                    f.put('\n');
                }
            };
            save("matrix_Z.txt", Zh, n, n);
            std::vector<std::complex<double>> v0(n), i0(n);
            for (int r = 0; r < n; ++r) {
                v0[r] = Bh[(size_t)r * ncols_];
                i0[r] = I_[(size_t)r * ncols_];
            }
            save("matrix_V.txt", v0, n, 1);
            save("martix_I.txt", i0, n, 1);
        }
    } catch (...) {
        release();
        throw;
    }
    release();
    std::cout << "success" << std::endl;
}

void DenseMoM::output() {
    std::cout << "Result Information\n";
    if (rtype_ == policy::BiRCS) {
        writeBistatic();
        timeLog("getBistaticRCS");
    } else {
        writeMonostatic();
        timeLog("getMonostaticRCS");
    }
    std::vector<component::CurrentData> cur;
    calculateSurfaceCurrent(&cur, 0);
    component::result::writeCurrent(dir_, cur);
    timeLog("getSurfaceCurrent");
}

void DenseMoM::clear() {
    if (X_) emfa_mat_destroy(X_);
    if (dmesh_) emfa_mesh_destroy(dmesh_);
    X_ = nullptr;
    dmesh_ = nullptr;
    mesh_.clear();
    ce_.clear();
}

// EMComponent/Result.h:48-104: phi cuts over theta, then the theta = 90 cut.
void DenseMoM::writeBistatic() {
    component::result::writeBistatic(dir_, scattering_, [&](int nd, const double *dirs, double *db) {
        check(emfa_rcs(ctx_, dmesh_, k_, X_, nd, dirs, nullptr, db), "RCS");
    });
}

// EMComponent/Result.h:107-155: one incidence per direction (backscatter).
void DenseMoM::writeMonostatic() {
    const auto &sca = scattering_;
    int nd = ncols_;
    std::vector<double> dirs(3 * (size_t)nd);
    std::vector<int> cols(nd);
    for (int j = 0; j < nd; ++j) {
        for (int c = 0; c < 3; ++c) dirs[3 * j + c] = rcs_dirs_[j][c];
        cols[j] = j;
    }
    std::vector<double> db(nd);
    check(emfa_rcs(ctx_, dmesh_, k_, X_, nd, dirs.data(), cols.data(), db.data()), "RCS");
    int phi_num = static_cast<int>((sca.phi_to - sca.phi_from) / sca.phi_delta) + 1;
    size_t o = 0;
    for (int i = 0; i < phi_num; ++i) {
        float phiDegree = sca.phi_from + i * sca.phi_delta;
        std::string str = std::to_string(phiDegree);
        std::ofstream out(dir_ + "/monostatic_rcs_phi_" + str.substr(0, str.find('.')) + ".dat");
        component::result::prefixRCSData(out);
        for (float thetaDegree = sca.theta_from; thetaDegree <= sca.theta_to; thetaDegree += sca.theta_delta)
            out << std::setw(10) << thetaDegree << '\t' << std::setw(10) << (float)db[o++] << '\n';
    }
    std::ofstream out(dir_ + "/monostatic_rcs_theta_90.dat");
    component::result::prefixRCSData(out);
    for (float phiDegree = 0; phiDegree < 360; phiDegree += 1)
        out << std::setw(10) << phiDegree << '\t' << std::setw(10) << (float)db[o++] << '\n';
}

// Per-triangle current magnitudes at the corners and the centroid,
// EMSolver/CFIE.cpp:727-794 (computed from the solved currents).
void DenseMoM::calculateSurfaceCurrent(std::vector<component::CurrentData> *currents, int col) const {
    const size_t nt = mesh_.getTriangleNum();
    const auto &xyz = mesh_.xyz();
    std::vector<std::vector<int>> tris(nt);
    for (int u = 0; u < unknowns_; ++u) {
        const auto &r = ce_.getRWGRef(u);
        tris[r.tpid].push_back(u);
        tris[r.tmid].push_back(u);
    }
    auto V = [&](int i) { return component::Vec3{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]}; };
    auto sub = [](component::Vec3 a, component::Vec3 b) { return component::Vec3{a.x - b.x, a.y - b.y, a.z - b.z}; };
    auto dist = [&](component::Vec3 a, component::Vec3 b) {
        auto d = sub(a, b);
        return std::sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
    };
    currents->resize(nt);
    // Corner magnitudes persist across triangles as in the reference
    // (CFIE.cpp:740 declares them outside the loop): at an open boundary a
    // triangle with fewer than three RWGs reports the previous triangle's
    // values at its other corners, and current.sc matches the reference's.
    double mg[3] = {0, 0, 0};
    for (size_t t = 0; t < nt; ++t) {
        auto &unks = tris[t];
        component::Vec3 vx[3];
        double len[3] = {0, 0, 0};
        for (size_t i = 0; i < unks.size() && i < 3; ++i) {
            const auto &r = ce_.getRWGRef(unks[i]);
            vx[i] = V(((int)t == r.tpid) ? r.vxp : r.vxm);
            len[i] = r.length;
        }
        if (unks.size() == 1) {
            const auto &r = ce_.getRWGRef(unks[0]);
            vx[1] = V(r.v1);
            vx[2] = V(r.v2);
        }
        if (unks.size() == 2) {
            const auto &r1 = ce_.getRWGRef(unks[0]);
            const auto &r2 = ce_.getRWGRef(unks[1]);
            vx[2] = (r1.v1 != r2.v1 && r1.v1 != r2.v2) ? V(r1.v2) : V(r1.v1);
        }
        component::Vec3 cen{(vx[0].x + vx[1].x + vx[2].x) / 3, (vx[0].y + vx[1].y + vx[2].y) / 3,
                            (vx[0].z + vx[1].z + vx[2].z) / 3};
        double la = dist(vx[0], vx[1]), lb = dist(vx[1], vx[2]), lc = dist(vx[2], vx[0]);
        double s = 0.5 * (la + lb + lc);
        double dom = 2 * std::sqrt(std::max(0.0, s * (s - la) * (s - lb) * (s - lc)));
        std::complex<double> cc[3] = {0, 0, 0};
        size_t m = std::min<size_t>(unks.size(), 3);
        for (size_t i = 0; i < m; ++i) {
            const auto &r = ce_.getRWGRef(unks[i]);
            double sg = ((int)t == r.tpid) ? 1.0 : -1.0;
            std::complex<double> f = sg * I_[(size_t)unks[i] * ncols_ + col] * (len[i] / dom);
            auto dv = sub(cen, vx[i]);
            cc[0] += f * dv.x;
            cc[1] += f * dv.y;
            cc[2] += f * dv.z;
            std::complex<double> nc[3] = {0, 0, 0};
            for (size_t j = 0; j < m; ++j) {
                if (i == j) continue;
                const auto &rj = ce_.getRWGRef(unks[j]);
                double sj = ((int)t == rj.tpid) ? 1.0 : -1.0;
                std::complex<double> fj = sj * I_[(size_t)unks[j] * ncols_ + col] * (len[j] / dom);
                auto d2 = sub(vx[i], vx[j]);
                nc[0] += fj * d2.x;
                nc[1] += fj * d2.y;
                nc[2] += fj * d2.z;
            }
            mg[i] = std::sqrt(std::norm(nc[0]) + std::norm(nc[1]) + std::norm(nc[2]));
        }
        auto &d = (*currents)[t];
        d.v1 = vx[0];
        d.v2 = vx[1];
        d.v3 = vx[2];
        d.magnc = std::sqrt(std::norm(cc[0]) + std::norm(cc[1]) + std::norm(cc[2]));
        d.magn1 = mg[0];
        d.magn2 = mg[1];
        d.magn3 = mg[2];
    }
}

// ------------------------------------------------------------------ CFIE
void CFIE::init(component::ConfigLoader *ploader) {
    alpha_ = ploader->getAlpha();
    DenseMoM::init(ploader);
}

// EMSolver/CFIE.cpp:35: "_CFIE_" + to_string(alpha).substr(0, 3)
std::string CFIE::folderTag() const { return "_CFIE_" + std::to_string((float)alpha_).substr(0, 3); }

// ------------------------------------------------------------------ MLFMA
void MLFMA::init(component::ConfigLoader *ploader) {
    alpha_ = ploader->getAlpha();
    iter_threshold_ = ploader->getDouble("iteration_threshold:", 1e-3);
    max_iter_ = ploader->getInt("max_iteration_number:", 1000);
    box_threshold_ = ploader->getDouble("mlfma_box_threshold:", 0.5);
    precond_requested_ = ploader->getString("preconditioning:", "false") == "true";
    if (precond_requested_) {
        // MLFMA.cpp:44-45 reads both thresholds (ConfigLoader.cpp:632-665)
        row_threshold_ = ploader->getDouble("row_threshold:", -1.0);
        col_threshold_ = ploader->getDouble("col_threshold:", -1.0);
        if (!(row_threshold_ > 0) || !(col_threshold_ > 0))
            throw component::ConfigInvalid("MLFMA: preconditioning needs positive row_threshold: and col_threshold:");
    }
    if (!(iter_threshold_ > 0) || max_iter_ < 1 || !(box_threshold_ > 0))
        throw component::ConfigInvalid("MLFMA: iteration_threshold, max_iteration_number and mlfma_box_threshold "
                                       "must be positive");
    DenseMoM::init(ploader);
    if (solver_ == LinearSolver::LU) throw component::ConfigInvalid("MLFMA: linear_solver must be BICGSTAB or GMRES");
}

// EMSolver/MLFMA.cpp:32: "_MLFMA_CFIE_" + to_string(alpha).substr(0, 3)
std::string MLFMA::folderTag() const { return "_MLFMA_CFIE_" + std::to_string((float)alpha_).substr(0, 3); }

void MLFMA::reportInfo(std::ostream &strm) const {
    DenseMoM::reportInfo(strm);
    strm << "MLFMA\n"
         << "  -> box threshold: " << box_threshold_ << " wavelengths\n"
         << "  -> " << (solver_ == LinearSolver::GMRES ? "GMRES(" + std::to_string(gmres_m_) + ")" : "BICGSTAB")
         << " to " << iter_threshold_ << ", at most " << max_iter_ << " iterations\n"
         << "  -> preconditioning: "
         << (precond_requested_ ? "sparse approximate inverse, row/col thresholds " + std::to_string(row_threshold_) +
                                      " / " + std::to_string(col_threshold_) + " leaf edges"
                                : std::string("off"))
         << '\n';
}

emfa_status MLFMA::createOperator(emfa_mlfma **op) {
    emfa_mlfma_params mp{EMFA_CFIE, k_, alpha_, box_threshold_, cfie_mode_};
    emfa_status st = emfa_mlfma_create(ctx_, dmesh_, &mp, op);
    if (st || !precond_requested_) return st;
    // MLFMA::solve calls calculatePreconditioner when preconditioning: true
    long long entries = 0;
    st = emfa_mlfma_precondition(ctx_, *op, dmesh_, row_threshold_, col_threshold_, &entries);
    if (st) {
        emfa_mlfma_destroy(*op);
        *op = nullptr;
        return st;
    }
    logger_ << "  -> Number of precondition: " << entries << '\n';
    return EMFA_OK;
}

void MLFMA::logOperator(emfa_mlfma *op) {
    emfa_mlfma_info info{};
    emfa_mlfma_get_info(op, &info);
    logger_ << "MLFMA\n  -> levels: " << info.levels << ", near entries: " << info.near_nnz
            << ", device bytes: " << info.bytes << '\n';
    for (int l = 2; l <= info.levels; ++l)
        logger_ << "  -> level " << l << ": " << info.boxes[l] << " boxes, L = " << info.multipoles[l] << '\n';
}

bool MLFMA::solveMatrixFree(emfa_mat *B) {
    const std::string name = operatorName();
    emfa_mlfma *op = nullptr;
    emfa_status st = createOperator(&op);
    if (st == EMFA_E_INVALID) throw std::runtime_error(name + ": " + emfa_last_error());
    check(st, (name + " setup").c_str());
    logOperator(op);
    timeLog(name + " setup");
    int iters = 0;
    double relres = 0;
    if (solver_ == LinearSolver::GMRES) {
        const int m = std::max(1, gmres_m_);
        emfa_gmres_report rep{};
        st = emfa_mlfma_gmres(ctx_, op, B, X_, m, iter_threshold_, (max_iter_ + m - 1) / m, &rep);
        iters = rep.iterations;
        relres = rep.relres;
    } else {
        // MLFMA::iterativeSolve (MLFMA.cpp:402-452), FMM::iterativeSolve (FMM.cpp:222-268)
        emfa_bicgstab_report rep{};
        st = emfa_mlfma_bicgstab(ctx_, op, B, X_, iter_threshold_, max_iter_, 0, &rep);
        iters = rep.iterations;
        relres = rep.relres;
    }
    emfa_mlfma_destroy(op);
    logger_ << "  -> Iterative number: " << iters << ", relative residual: " << relres << '\n';
    std::cout << "  -> Iterative number: " << iters << std::endl;
    if (st == EMFA_E_NOCONV)
        logger_ << "  -> WARNING: Iteration cannot converge\n";
    else
        check(st, solverName(solver_));
    return true;
}

// ------------------------------------------------------------------ FMM
void FMM::init(component::ConfigLoader *ploader) {
    alpha_ = ploader->getAlpha();
    iter_threshold_ = ploader->getDouble("iteration_threshold:", 1e-3);
    max_iter_ = ploader->getInt("max_iteration_number:", 1000);
    // FMM.cpp:43 reads fmm_box_length: unconditionally (ConfigLoader.cpp:536-541)
    box_threshold_ = ploader->getDouble("fmm_box_length:", -1.0);
    if (!(box_threshold_ > 0)) throw component::ConfigInvalid("FMM: fmm_box_length: must be given and positive");
    if (!(iter_threshold_ > 0) || max_iter_ < 1)
        throw component::ConfigInvalid("FMM: iteration_threshold and max_iteration_number must be positive");
    DenseMoM::init(ploader);
    if (solver_ == LinearSolver::LU) throw component::ConfigInvalid("FMM: linear_solver must be BICGSTAB or GMRES");
}

// EMSolver/FMM.cpp:33: "_FMM_CFIE_" + to_string(alpha).substr(0, 3)
std::string FMM::folderTag() const { return "_FMM_CFIE_" + std::to_string((float)alpha_).substr(0, 3); }

void FMM::reportInfo(std::ostream &strm) const {
    DenseMoM::reportInfo(strm);
    strm << "FMM\n"
         << "  -> box length: " << box_threshold_ << " wavelengths\n"
         << "  -> " << (solver_ == LinearSolver::GMRES ? "GMRES(" + std::to_string(gmres_m_) + ")" : "BICGSTAB")
         << " to " << iter_threshold_ << ", at most " << max_iter_ << " iterations\n";
}

emfa_status FMM::createOperator(emfa_mlfma **op) {
    emfa_mlfma_params mp{EMFA_CFIE, k_, alpha_, box_threshold_, cfie_mode_};
    return emfa_fmm_create(ctx_, dmesh_, &mp, op);
}

// FMM::reportFFMInfo / reportBoxInfo (FMM.cpp:729-772): the operator's sizes in
// the log and one line per box in ffm_box_log.txt
void FMM::logOperator(emfa_mlfma *op) {
    emfa_mlfma_info info{};
    emfa_mlfma_get_info(op, &info);
    const int nb = info.boxes[1], L = info.multipoles[1];
    const long long K = 2LL * L * L, n = unknowns_;
    std::vector<int> nn(nb), nf(nb), nu(nb);
    emfa_mlfma_boxes(op, 1, nn.data(), nf.data(), nu.data());
    const size_t mb = 1024 * 1024;
    const size_t near_mem = (size_t)info.near_nnz * 2 * sizeof(float) / mb;
    const size_t rdcv_mem = 12 * (size_t)n * K * sizeof(float) / mb;
    const size_t tran_mem = 2 * (size_t)info.far_pairs[1] * K * sizeof(float) / mb;
    const size_t orig_mem = 2 * sizeof(float) * (size_t)n * n / mb;
    const size_t total_mem = near_mem + rdcv_mem + tran_mem;
    auto info_lines = [&](std::ostream &o) {
        o << "FMM Information:\n"
          << "  -> Total boxes: " << (long long)info.grid[0] * info.grid[1] * info.grid[2] << '\n'
          << "  -> Valid boxes: " << nb << '\n'
          << "  -> Average per box: " << (nb ? n / nb : 0) << '\n'
          << "  -> Number of near: " << info.near_nnz << '\n'
          << "  -> Sparse rate: " << 100.0f * (float)info.near_nnz / (float)n / (float)n << "%\n"
          << "  -> Number of multipoles: " << L << '\n'
          << "  -> Near impedance: " << near_mem << " MB\n"
          << "  -> Transfer (per box pair, as the reference stores it): " << tran_mem << " MB\n"
          << "  -> Rad and Recv: " << rdcv_mem << " MB\n"
          << "  -> Total memory" << total_mem << " MB\n"
          << "  -> Saved memory: " << (total_mem > orig_mem ? "-" : "")
          << (total_mem > orig_mem ? total_mem - orig_mem : orig_mem - total_mem) << " MB\n"
          << "  -> Grid: " << info.grid[0] << 'x' << info.grid[1] << 'x' << info.grid[2] << " boxes of "
          << info.box_edge[1] << " m, device bytes: " << info.bytes << '\n';
    };
    info_lines(logger_);
    std::ofstream box_log(dir_ + "/ffm_box_log.txt");
    info_lines(box_log);
    char line[128];
    for (int b = 0; b < nb; ++b) {
        std::snprintf(line, sizeof(line), "    -> Box %4d:    near: %3d    far: %5d    unknowns: %5d\n", b + 1, nn[b],
                      nf[b], nu[b]);
        box_log << line;
    }
}

// ------------------------------------------------------------------ FastWT
FastWT::~FastWT() { clear(); }

void FastWT::check(emfa_status st, const char *what) const {
    if (st == EMFA_OK) return;
    std::string msg = std::string(what) + ": " + emfa_last_error();
    if (st == EMFA_E_INVALID) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

void FastWT::init(component::ConfigLoader *ploader) {
    EM::init(ploader);
    if (rtype_ == policy::Rad)
        throw component::ConfigInvalid("result_type Radiation (delta-gap .rad excitation) is not part of this build");
    alpha_ = ploader->getAlpha();
    // FastWT.cpp:35: "_FastWT_" + to_string(alpha).substr(0, 3)
    dir_ = tool::creatFolder(dir_, folder_name_ + "_FastWT_" + std::to_string((float)alpha_).substr(0, 3));
    mesh_.loadMeshFile(ploader->getMeshPath());
    ce_.buildCommonEdge(mesh_);
    const auto &xyz = mesh_.xyz();
    const auto &tri = mesh_.tri();
    const int nv = (int)(xyz.size() / 3), nt = (int)(tri.size() / 3);
    int n = 0, n_nc = 0, n_lines = 0;
    check(emfa_contour_build(nv, xyz.data(), nt, tri.data(), nullptr, nullptr, 0, &n, nullptr, nullptr, nullptr, 0,
                             &n_nc, &n_lines),
          "contour edges");
    if (n <= 0) throw component::FileError("mesh has no half-RWG unknowns");
    hrwg_.assign(6 * (size_t)n, 0);
    hlen_.assign(n, 0.0);
    line_ptr_.assign(n_nc + 1, 0);
    line_tri_.assign(std::max(1, n_lines), 0);
    line_seg_.assign(6 * (size_t)std::max(1, n_lines), 0.0);
    check(emfa_contour_build(nv, xyz.data(), nt, tri.data(), hrwg_.data(), hlen_.data(), n, &n, line_ptr_.data(),
                             line_tri_.data(), line_seg_.data(), (int)line_tri_.size(), &n_nc, &n_lines),
          "contour edges");
    unknowns_ = n;
    n_nc_ = n_nc;
    iter_threshold_ = ploader->getDouble("iteration_threshold:", 1e-3);
    max_iter_ = ploader->getInt("max_iteration_number:", 1000);
    precond_ = ploader->getString("preconditioning:", "false") == "true";
    if (precond_) {
        row_threshold_ = ploader->getDouble("row_threshold:", -1.0);
        col_threshold_ = ploader->getDouble("col_threshold:", -1.0);
        if (!(row_threshold_ > 0) || !(col_threshold_ > 0))
            throw component::ConfigInvalid("FastWT: preconditioning needs positive row_threshold: and col_threshold:");
    }
    // FastWT.cpp:48: beta = stabilization_factor * |log10(average_size)|
    const double factor = ploader->getDouble("stabilization_factor:", -1.0);
    const double avg = ploader->getDouble("average_size:", -1.0);
    if (!(factor >= 0) || !(avg > 0))
        throw component::ConfigInvalid("FastWT: stabilization_factor: (>= 0) and average_size: (> 0) must be given");
    beta_ = factor * std::fabs(std::log10(avg));
    box_threshold_ = ploader->getDouble("mlfma_box_threshold:", 0.5);
    if (!(iter_threshold_ > 0) || max_iter_ < 1 || !(box_threshold_ > 0))
        throw component::ConfigInvalid("FastWT: iteration_threshold, max_iteration_number and mlfma_box_threshold "
                                       "must be positive");
    std::string ls = ploader->getString("linear_solver:", "BICGSTAB");
    if (ls != "BICGSTAB" && ls != "GMRES") throw component::ConfigInvalid("FastWT: linear_solver must be BICGSTAB or GMRES");
    gmres_ = ls == "GMRES";
    gmres_m_ = ploader->getInt("gmres_restart:", 30);
    std::string cm = ploader->getString("cfie_self_term:", "reference");
    if (cm != "reference" && cm != "accurate")
        throw component::ConfigInvalid("cfie_self_term must be reference or accurate");
    cfie_mode_ = cm == "accurate" ? 1 : 0;
    device_ = ploader->getInt("device:", 0);
    k_ = 2.0 * kPi * incidence_.freq / kC0;
    logger_.open(dir_ + "/sca_parallel_runtime.log");
    logger_ << ploader->getReport();
    reportInfo(logger_);
    timeLog("init");
}

void FastWT::reportInfo(std::ostream &strm) const {
    mesh_.reportInfo(strm);
    int conformal = 0;
    for (int u = 0; u < unknowns_; ++u) conformal += hrwg_[6 * (size_t)u + 5] == 0;
    // ContourEdge::reportInfo (ContourEdge.cpp:103-117)
    strm << "ContourEdge Information\n"
         << "  -> Total half-RWGs:" << unknowns_ << '\n'
         << "     conformal: " << conformal << '\n'
         << "     nonconformal: " << n_nc_ << '\n'
         << "FastWT Information\n"
         << "  -> Stabilization parameter:" << beta_ << '\n'
         << "  -> box threshold: " << box_threshold_ << " wavelengths\n"
         << "  -> " << (gmres_ ? "GMRES(" + std::to_string(gmres_m_) + ")" : std::string("BICGSTAB")) << " to "
         << iter_threshold_ << ", at most " << max_iter_ << " iterations\n"
         << "  -> preconditioning: "
         << (precond_ ? "sparse approximate inverse, row/col thresholds " + std::to_string(row_threshold_) + " / " +
                            std::to_string(col_threshold_) + " leaf edges"
                      : std::string("off"))
         << '\n';
}

// FastWT::reportLevelDetail (FastWT.cpp:1295-1318): per level and box its
// near and far boxes and unknowns (the reference's child count is not kept by
// the device operator)
void FastWT::reportLevelDetail(emfa_mlfma *op) const {
    emfa_mlfma_info info{};
    emfa_mlfma_get_info(op, &info);
    std::ofstream f(dir_ + "/level_detail_info.txt");
    f << "MLFMA Level Information\n  -> Number of level: " << info.levels + 1 << "\n  -> Number of valid boxes:\n";
    for (int l = 1; l <= info.levels; ++l) f << "    => level " << std::setw(2) << l << std::setw(15) << ':' << info.boxes[l] << '\n';
    f << "Detail of level box: \n";
    for (int l = 1; l <= info.levels; ++l) {
        std::vector<int> nn(info.boxes[l]), nf(info.boxes[l]), nu(info.boxes[l]);
        emfa_mlfma_boxes(op, l, nn.data(), nf.data(), nu.data());
        f << "  -> level " << l << " :\n";
        for (int b = 0; b < info.boxes[l]; ++b)
            f << "    => box " << std::setw(8) << b + 1 << ":    near: " << std::setw(3) << nn[b]
              << "    far: " << std::setw(3) << nf[b] << "    unknown: " << nu[b] << '\n';
    }
}

void FastWT::solve() {
    std::cout << "Solve Information\n";
    ctx_ = device_context(device_);
    const auto &xyz = mesh_.xyz();
    const auto &tri = mesh_.tri();
    if (ce_.getCommonEdgeNum() <= 0) throw component::FileError("FastWT: the mesh has no shared edge");
    check(emfa_mesh_upload(ctx_, xyz.data(), (int)(xyz.size() / 3), tri.data(), (int)(tri.size() / 3),
                           ce_.flat().data(), ce_.getCommonEdgeNum(), &dmesh_),
          "mesh upload");
    const int n = unknowns_;
    emfa_mat *B = nullptr;
    try {
        emfa_mlfma_params mp{EMFA_CFIE, k_, alpha_, box_threshold_, cfie_mode_};
        check(emfa_fastwt_create(ctx_, dmesh_, n, hrwg_.data(), hlen_.data(), n_nc_, line_ptr_.data(),
                                 line_tri_.data(), line_seg_.data(), &mp, beta_, &op_),
              "FastWT setup");
        emfa_mlfma_info info{};
        emfa_mlfma_get_info(op_, &info);
        logger_ << "FastWT operator\n  -> levels: " << info.levels << ", near entries: " << info.near_nnz
                << ", device bytes: " << info.bytes << '\n';
        for (int l = 2; l <= info.levels; ++l)
            logger_ << "  -> level " << l << ": " << info.boxes[l] << " boxes, L = " << info.multipoles[l] << '\n';
        reportLevelDetail(op_);
        timeLog("FastWT setup");
        if (precond_) {
            long long entries = 0;
            check(emfa_mlfma_precondition(ctx_, op_, dmesh_, row_threshold_, col_threshold_, &entries),
                  "Preconditioner");
            logger_ << "  -> Number of precondition: " << entries << '\n';
            timeLog("Preconditioner");
        }
        std::vector<double> inc(9);
        for (int c = 0; c < 3; ++c) {
            inc[c] = inc_k_[c];
            inc[3 + c] = inc_e_[c];
            inc[6 + c] = inc_h_[c];
        }
        check(emfa_mat_create(ctx_, n, 1, EMFA_F64, &B), "excitation");
        check(emfa_fastwt_rhs(ctx_, op_, dmesh_, B, 1, inc.data()), "fillVbyBox");
        timeLog("fillVbyBox");
        check(emfa_mat_create(ctx_, n, 1, EMFA_F64, &X_), "solution");
        int iters = 0;
        double relres = 0;
        emfa_status st;
        if (gmres_) {
            const int m = std::max(1, gmres_m_);
            emfa_gmres_report rep{};
            st = emfa_mlfma_gmres(ctx_, op_, B, X_, m, iter_threshold_, (max_iter_ + m - 1) / m, &rep);
            iters = rep.iterations;
            relres = rep.relres;
        } else {
            emfa_bicgstab_report rep{};
            st = emfa_mlfma_bicgstab(ctx_, op_, B, X_, iter_threshold_, max_iter_, 0, &rep);
            iters = rep.iterations;
            relres = rep.relres;
        }
        if (st == EMFA_E_NOCONV)
            logger_ << "  -> WARNING: Iteration cannot converge\n";
        else
            check(st, gmres_ ? "GMRES" : "BICGSTAB");
        logger_ << "  -> Iterative number: " << iters << ", relative residual: " << relres << '\n';
        std::cout << "  -> Iterative number: " << iters << std::endl;
        I_.assign(n, {0, 0});
        check(emfa_mat_download(X_, reinterpret_cast<double *>(I_.data())), "solution download");
        timeLog("iterativeSolve");
    } catch (...) {
        if (B) emfa_mat_destroy(B);
        throw;
    }
    emfa_mat_destroy(B);
    std::cout << "success" << std::endl;
}

void FastWT::output() {
    std::cout << "Result Information\n";
    // FastWT::output (FastWT.cpp:107-124): bistatic RCS, then the current
    component::result::writeBistatic(dir_, scattering_, [&](int nd, const double *dirs, double *db) {
        check(emfa_fastwt_rcs(ctx_, op_, dmesh_, X_, 0, nd, dirs, db), "RCS");
    });
    timeLog("getBistaticRCS");
    std::vector<component::CurrentData> cur;
    calculateSurfaceCurrent(&cur);
    component::result::writeCurrent(dir_, cur);
    timeLog("getSurfaceCurrent");
}

// FastWT::calculateSurfaceCurrent (FastWT.cpp:1426-1472): per triangle, its
// half-RWGs in list order (at most three), centre and corner magnitudes from
// l / (2 A) (r - v_free).  The corner vertices, lengths and magnitudes persist
// from the previous triangle where a triangle has fewer than three half-RWGs,
// as the reference declares them outside its loop.
void FastWT::calculateSurfaceCurrent(std::vector<component::CurrentData> *currents) const {
    const size_t nt = mesh_.getTriangleNum();
    const auto &xyz = mesh_.xyz();
    std::vector<std::vector<int>> tris(nt);
    for (int u = 0; u < unknowns_; ++u) tris[hrwg_[6 * (size_t)u + 3]].push_back(u);
    auto V = [&](int i) { return component::Vec3{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]}; };
    auto sub = [](component::Vec3 a, component::Vec3 b) { return component::Vec3{a.x - b.x, a.y - b.y, a.z - b.z}; };
    auto dist = [&](component::Vec3 a, component::Vec3 b) {
        auto d = sub(a, b);
        return std::sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
    };
    currents->resize(nt);
    component::Vec3 vx[3];
    double len[3] = {0, 0, 0}, mg[3] = {0, 0, 0};
    for (size_t t = 0; t < nt; ++t) {
        const auto &unks = tris[t];
        const size_t m = std::min<size_t>(unks.size(), 3);
        for (size_t i = 0; i < m; ++i) {
            vx[i] = V(hrwg_[6 * (size_t)unks[i] + 2]);
            len[i] = hlen_[unks[i]];
        }
        const component::Vec3 cen{(vx[0].x + vx[1].x + vx[2].x) / 3, (vx[0].y + vx[1].y + vx[2].y) / 3,
                                  (vx[0].z + vx[1].z + vx[2].z) / 3};
        const double la = dist(vx[0], vx[1]), lb = dist(vx[1], vx[2]), lc = dist(vx[2], vx[0]);
        const double s = 0.5 * (la + lb + lc);
        const double dom = 2 * std::sqrt(std::max(0.0, s * (s - la) * (s - lb) * (s - lc)));
        std::complex<double> cc[3] = {0, 0, 0};
        for (size_t i = 0; i < m; ++i) {
            const std::complex<double> f = I_[unks[i]] * (len[i] / dom);
            const auto dv = sub(cen, vx[i]);
            cc[0] += f * dv.x;
            cc[1] += f * dv.y;
            cc[2] += f * dv.z;
            std::complex<double> nc[3] = {0, 0, 0};
            for (size_t j = 0; j < m; ++j) {
                if (i == j) continue;
                const std::complex<double> fj = I_[unks[j]] * (len[j] / dom);
                const auto d2 = sub(vx[i], vx[j]);
                nc[0] += fj * d2.x;
                nc[1] += fj * d2.y;
                nc[2] += fj * d2.z;
            }
            mg[i] = std::sqrt(std::norm(nc[0]) + std::norm(nc[1]) + std::norm(nc[2]));
        }
        auto &d = (*currents)[t];
        d.v1 = vx[0];
        d.v2 = vx[1];
        d.v3 = vx[2];
        d.magnc = std::sqrt(std::norm(cc[0]) + std::norm(cc[1]) + std::norm(cc[2]));
        d.magn1 = mg[0];
        d.magn2 = mg[1];
        d.magn3 = mg[2];
    }
}

void FastWT::clear() {
    if (X_) emfa_mat_destroy(X_);
    if (op_) emfa_mlfma_destroy(op_);
    if (dmesh_) emfa_mesh_destroy(dmesh_);
    X_ = nullptr;
    op_ = nullptr;
    dmesh_ = nullptr;
    mesh_.clear();
    ce_.clear();
}

}  // namespace mom

namespace solver {

void loadSolver(const std::string &conf_file) {
    component::ConfigLoader loader;
    loader.load(conf_file);
    auto solver_type = loader.getSolverType();
    if (solver_type == "EFIE")
        invoke_solver<mom::EFIE>(&loader);
    else if (solver_type == "CFIE")
        invoke_solver<mom::CFIE>(&loader);
    else if (solver_type == "MFIE")
        invoke_solver<mom::MFIE>(&loader);
    else if (solver_type == "MLFMA")
        invoke_solver<mom::MLFMA>(&loader);
    else if (solver_type == "FMM")
        invoke_solver<mom::FMM>(&loader);
    else if (solver_type == "FastWT")
        invoke_solver<mom::FastWT>(&loader);
    else
        throw component::ConfigInvalid("undefined EM solver type (" + solver_type + ')');
}

}  // namespace solver
