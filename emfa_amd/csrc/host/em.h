// Solver lifecycle of the drop-in: mom::EMBase / mom::EM with the five
// virtuals of EMSolver/EM.h:28-32, and the dense MoM solvers EFIE, CFIE and
// MFIE whose fill, excitation, factorisation, solve and far field run on the
// gfx950 device through the C ABI of include/emfa_hip.h.
#pragma once
#include <array>
#include <complex>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "core.h"
#include "emfa_hip.h"

namespace mom {

// EMSolver/EMBase.h:10-17
class EMBase {
  public:
    EMBase() {}
    virtual ~EMBase() {}
    EMBase(const EMBase &) = delete;
    EMBase &operator=(const EMBase &) = delete;
};

// EMSolver/EM.h:21-61 and EM.cpp:24-61
class EM : public EMBase {
  public:
    EM();
    ~EM() override;
    virtual void init(component::ConfigLoader *ploader);
    virtual void solve() {}
    virtual void output() {}
    virtual void clear() {}
    virtual void reportInfo(std::ostream &strm) const { (void)strm; }

    const std::string &dir() const { return dir_; }

  protected:
    void timeLog(const std::string &what);
    policy::ResultType rtype_ = policy::BiRCS;
    std::string dir_, folder_name_;
    std::ofstream logger_;
    std::ostringstream time_logger_;
    tool::Clock clock_;
    component::Incidence incidence_;
    component::Scattering scattering_;
    component::Radiation radiation_;
    component::Nearfield nearfield_;
    std::array<double, 3> inc_k_{}, inc_e_{}, inc_h_{};
    double lamda = 0;
};

// Device context shared by all solvers of the process (created lazily).
emfa_ctx *device_context(int device);

// Dense method of moments on the device.  Build-specific configuration keys
// (absent from the reference's CONFIG.yml, defaults in brackets):
//   precision:        FP64 | FP32          [FP64]  FP32 = complex-FP32 fill and
//                                                  LU with FP64 iterative refinement
//   quadrature_order: 3 | 7                [3]     triangle rule of the fill
//   linear_solver:    LU | GMRES | BICGSTAB [LU; CFIE: BICGSTAB as CFIE.cpp:76;
//                                                  MLFMA: BICGSTAB as MLFMA.cpp:402-452]
//   gmres_restart:    m                    [30]
//   gmres_tolerance:  relative residual    [1e-6]
//   bicgstab_tolerance:      ||s|| / ||b|| [0.01, CFIE.cpp:76]
//   bicgstab_max_iterations: steps         [500, CFIE.cpp:76]
//   cfie_self_term:   reference | accurate [reference]
//   write_matrix:     true | false         [false] ASCII Z/V/I as EFIE.cpp:932-950
//   device:           HIP ordinal          [0]
class DenseMoM : public EM {
  public:
    ~DenseMoM() override;
    void init(component::ConfigLoader *ploader) override;
    void solve() override;
    void output() override;
    void clear() override;
    void reportInfo(std::ostream &strm) const override;

    // results, for callers that drive the solver directly
    int unknowns() const { return unknowns_; }
    const std::vector<std::complex<double>> &current() const { return I_; }  // n x ncols, row-major
    int currentColumns() const { return ncols_; }
    double wavenumber() const { return k_; }

  protected:
    virtual int fillKind() const = 0;  // EMFA_EFIE or EMFA_CFIE
    virtual int rhsKind() const = 0;   // EMFA_EFIE, EMFA_CFIE or EMFA_MFIE
    virtual double alpha() const { return 1.0; }
    virtual std::string folderTag() const = 0;
    // the reference's solver for this formulation when linear_solver: is absent
    virtual std::string defaultSolver() const { return "LU"; }

    // Matrix-free solvers override this: solve Z X_ = B without a dense Z
    // and return true.  The dense path (fill, then LU or GMRES) runs otherwise.
    virtual bool solveMatrixFree(emfa_mat *B) {
        (void)B;
        return false;
    }
    void check(emfa_status st, const char *what) const;
    void calculateSurfaceCurrent(std::vector<component::CurrentData> *currents, int col) const;
    void writeBistatic();
    void writeMonostatic();
    void writeMatrices();

    component::Mesh mesh_;
    component::CommonEdge ce_;
    int unknowns_ = 0;
    double k_ = 0;
    std::complex<double> sigma_{0, 0};
    enum class LinearSolver { LU, GMRES, BICGSTAB };
    int prec_ = EMFA_F64, quad_ = 3, cfie_mode_ = 0, gmres_m_ = 30, device_ = 0, bicg_max_ = 500;
    double gmres_tol_ = 1e-6, bicg_tol_ = 0.01;
    LinearSolver solver_ = LinearSolver::LU;
    bool write_matrix_ = false;
    static const char *solverName(LinearSolver s);
    std::vector<std::array<double, 3>> rhs_angles_;  // (theta, phi, pole) per column
    std::vector<std::array<double, 3>> rcs_dirs_;    // monostatic: direction per column
    int ncols_ = 0;
    std::vector<std::complex<double>> I_;

    emfa_ctx *ctx_ = nullptr;
    emfa_mesh *dmesh_ = nullptr;
    emfa_mat *X_ = nullptr;  // solution (F64), kept for the far field
};

class EFIE : public DenseMoM {
  protected:
    int fillKind() const override { return EMFA_EFIE; }
    int rhsKind() const override { return EMFA_EFIE; }
    std::string folderTag() const override { return "_EFIE"; }
};

class CFIE : public DenseMoM {
  public:
    void init(component::ConfigLoader *ploader) override;

  protected:
    int fillKind() const override { return EMFA_CFIE; }
    int rhsKind() const override { return EMFA_CFIE; }
    double alpha() const override { return alpha_; }
    std::string folderTag() const override;
    std::string defaultSolver() const override { return "BICGSTAB"; }
    double alpha_ = 0.5;
};

// MFIE: the magnetic half of CFIE (alpha = 0), solved with the dense LU as
// EMSolver/MFIE.cpp:50-73 does (arma::solve).
class MFIE : public DenseMoM {
  protected:
    int fillKind() const override { return EMFA_CFIE; }
    int rhsKind() const override { return EMFA_MFIE; }
    double alpha() const override { return 0.0; }
    std::string folderTag() const override { return "_MFIE"; }
};

// EMSolver/MLFMA.cpp: the CFIE operator applied by the multilevel fast
// multipole method (near matrix + aggregation / translation /
// disaggregation on the device), solved by BiCGSTAB to iteration_threshold
// in at most max_iteration_number steps as MLFMA.cpp:402-452 does, or by
// GMRES(gmres_restart) with linear_solver: GMRES.  Keys: alpha:,
// iteration_threshold:, max_iteration_number:, mlfma_box_threshold:;
// preconditioning: is read and reported but the sparse approximate inverse
// is not part of this build.
class MLFMA : public DenseMoM {
  public:
    void init(component::ConfigLoader *ploader) override;
    void reportInfo(std::ostream &strm) const override;

  protected:
    int fillKind() const override { return EMFA_CFIE; }
    int rhsKind() const override { return EMFA_CFIE; }
    double alpha() const override { return alpha_; }
    std::string folderTag() const override;
    bool solveMatrixFree(emfa_mat *B) override;
    std::string defaultSolver() const override { return "BICGSTAB"; }
    // the fast operator of this solver and what its set-up reports
    virtual emfa_status createOperator(emfa_mlfma **op);
    virtual void logOperator(emfa_mlfma *op);
    virtual const char *operatorName() const { return "MLFMA"; }
    double alpha_ = 0.5, box_threshold_ = 0.5, iter_threshold_ = 1e-3;
    double row_threshold_ = 0.9, col_threshold_ = 0.87;  // leaf edges (preconditioner)
    int max_iter_ = 1000;
    bool precond_requested_ = false;
};

// Single-level FMM (EMSolver/FMM.cpp): CFIE on a uniform box grid of
// `fmm_box_length:` wavelengths, solved with BiCGSTAB (FMM.cpp:222-268).
class FMM : public MLFMA {
  public:
    void init(component::ConfigLoader *ploader) override;
    void reportInfo(std::ostream &strm) const override;

  protected:
    std::string folderTag() const override;
    emfa_status createOperator(emfa_mlfma **op) override;
    void logOperator(emfa_mlfma *op) override;
    const char *operatorName() const override { return "FMM"; }
};

// EMSolver/FastWT.cpp: the IP-SIE (interior-penalty surface integral
// equation) fast solver on half-RWGs.  The half-RWG list with its contour
// lines is emfa_contour_build's (ContourEdge with policy::HALF_RWG); the
// operator, excitation and far field run on the device (emfa_fastwt_*), the
// solve is BiCGSTAB (FastWT::iterativeSolve, FastWT.cpp:916-962) to
// iteration_threshold: in at most max_iteration_number: steps, or
// GMRES(gmres_restart) with linear_solver: GMRES; preconditioning: true
// applies the sparse approximate inverse (FastWT.cpp:808-914).  Keys as
// FastWT::init (FastWT.cpp:26-70): alpha:, mlfma_box_threshold:,
// stabilization_factor: and average_size: (beta = factor |log10(size)|),
// row_threshold: / col_threshold: with preconditioning.  Output folder
// <mesh>_FastWT_<alpha>, log sca_parallel_runtime.log, the bistatic RCS files
// and current.sc of FastWT::output (FastWT.cpp:107-124).
class FastWT : public EM {
  public:
    ~FastWT() override;
    void init(component::ConfigLoader *ploader) override;
    void solve() override;
    void output() override;
    void clear() override;
    void reportInfo(std::ostream &strm) const override;

    int unknowns() const { return unknowns_; }
    const std::vector<std::complex<double>> &current() const { return I_; }  // half-RWG order
    const std::vector<int> &halfRWGs() const { return hrwg_; }             // n x 6
    double wavenumber() const { return k_; }

  protected:
    void check(emfa_status st, const char *what) const;
    void calculateSurfaceCurrent(std::vector<component::CurrentData> *currents) const;
    void reportLevelDetail(emfa_mlfma *op) const;

    component::Mesh mesh_;
    component::CommonEdge ce_;  // RWG list of the mesh (the device mesh needs one)
    std::vector<int> hrwg_, line_ptr_, line_tri_;
    std::vector<double> hlen_, line_seg_;
    int unknowns_ = 0, n_nc_ = 0, max_iter_ = 1000, gmres_m_ = 30, cfie_mode_ = 0, device_ = 0;
    double k_ = 0, alpha_ = 0.5, beta_ = 0, box_threshold_ = 0.5, iter_threshold_ = 1e-3;
    double row_threshold_ = 0.9, col_threshold_ = 0.87;
    bool precond_ = false, gmres_ = false;
    std::vector<std::complex<double>> I_;
    emfa_ctx *ctx_ = nullptr;
    emfa_mesh *dmesh_ = nullptr;
    emfa_mlfma *op_ = nullptr;
    emfa_mat *X_ = nullptr;
};

}  // namespace mom

namespace solver {

// EMSolver/Solver.h:21-29
template <typename QSolver>
void invoke_solver(component::ConfigLoader *ploader) {
    auto s = std::make_unique<QSolver>();
    s->init(ploader);
    s->reportInfo(std::cout);
    s->solve();
    s->output();
    s->clear();
}

// EMSolver/Solver.cpp:45-95: string registry on "EMSolver:"
void loadSolver(const std::string &conf_file);

}  // namespace solver
