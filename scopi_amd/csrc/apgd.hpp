// apgd.hpp -- host side of the device optim slot (B assembly + projected gradient).
#pragma once

#include "common.hpp"
#include "sort_scan.hpp"

namespace scopi_hip {

// Scalar state of the projected-gradient loop, owned by the device.  The last
// block of the row kernel runs the reference's control logic
// (include/scopi/solvers/apgd.hpp:164-213) on it, so the host only polls `done`.
struct ApgdCtrl {
    double L, alpha, theta, beta;
    double L_acc;        // L of the last accepted trial (before the 0.97 decay)
    double tol, rtol, dt2;
    double res;          // last ||lambda_{k+1} - lambda_k||
    double red[9];       // sums of the trial kernel (0..6, 8) and of the row kernel (7)
    int64_t iter, backtracks, max_ite, q_applies, consec_bt;
    int32_t cur, prev, trial;
    int32_t done, status, method, dynamic, low_precision;
    uint32_t count_trial, count_row;
};

// Coefficients of the contact Jacobian B (one block row per contact).
// NoFriction: row c = [ n_c | a_i,c ] on particle i and -[ n_c | a_j,c ] on j,
//   a_k,c = R_k^T (r_k x n_c), r_k = p_k - x_k.
// Friction: rows (c, d), d = 0..2 = [ e_d | K_i,c[d] ] on i, -[ e_d | K_j,c[d] ] on j,
//   K_k,c[d] = R_k^T (r_k x e_d).
template <class T>
struct BMatrix {
    int64_t cap = 0;
    T* coef = nullptr;   // NoFriction: 9 x cap (n, a_i, a_j); Friction: 21 x cap (n, K_i, K_j)
    T* c = nullptr;      // rows: cap (NoFriction) or 3 cap (Friction), component-major
    T* mu = nullptr;     // friction coefficient per contact (Friction)
    T* minv = nullptr;   // 6 x pcap
    T* lam = nullptr;    // 3 slots x rows
    T* qlam = nullptr;   // 3 slots x rows
    T* w = nullptr;      // 6 x pcap
};

class Solver {
  public:
    ~Solver();
    // Assembles B and c from the device contacts and the a-priori velocities,
    // solves, and leaves u = v + dt M^-1 B^T lambda in `u_out` / `w_out` (3 x pcap each).
    int solve(const Particles& P, const Contacts& C, int64_t nc, const scopi_hip_params& prm, double dt,
              hipStream_t s, scopi_hip_solve_info* info);
    // lambda of the last solve (FP64, row layout of the reference: nc or 3 nc).
    void download_lambda(double* host, hipStream_t s) const;
    const double* u() const { return u_; }
    const double* w() const { return w_; }
    int64_t pcap() const { return pcap_; }
    int64_t last_rows() const { return rows_; }
    int64_t last_nc() const { return nc_; }
    bool has_solution() const { return has_solution_; }

    void set_timing(bool on) { timing_ = on; }
    void timing(double* out) const;

  private:
    void reserve(int64_t na, int64_t nc, bool friction, bool fp32);
    template <class T>
    void run_phase(BMatrix<T>& B, const Particles& P, const Contacts& C, int64_t na, int64_t nc, bool friction,
                   const scopi_hip_params& prm, double dt, double tol, double rtol, int64_t max_ite, bool warm,
                   bool keep_L, hipStream_t s, scopi_hip_solve_info& info);

    RadixSorter sorter_;
    Scanner scanner_;
    int64_t pcap_ = 0, ccap_ = 0;
    BMatrix<double> b64_;
    BMatrix<float> b32_;
    bool b32_alloc_ = false;
    // particle-major transpose: contacts where the particle is i (contiguous)
    // and where it is j (stable-sorted list of contact indices)
    int* istart_ = nullptr;
    int* iend_ = nullptr;
    int* jstart_ = nullptr;
    int* jend_ = nullptr;
    uint32_t* jkeys_ = nullptr;
    int32_t* jlist_ = nullptr;
    double* u_ = nullptr;  // 3 x pcap
    double* w_ = nullptr;  // 3 x pcap
    ApgdCtrl* ctrl_ = nullptr;
    ApgdCtrl* h_ctrl_ = nullptr;  // pinned
    double* partial_ = nullptr;
    int64_t partial_cap_ = 0;
    int64_t rows_ = 0, nc_ = 0;
    int final_slot_ = 0;
    bool has_solution_ = false;
    bool last_friction_ = false;
    // kernel timing (HIP events on the solve stream)
    bool timing_ = false;
    double t_bt_ms_ = 0, t_row_ms_ = 0, t_trial_ms_ = 0;
    int64_t n_bt_ = 0, n_row_ = 0, n_trial_ = 0;
};

}  // namespace scopi_hip
