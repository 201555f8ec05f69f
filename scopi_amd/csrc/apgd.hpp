// apgd.hpp -- host side of the device optim slot (B assembly + projected gradient).
#pragma once

#include "common.hpp"
#include "sort_scan.hpp"

#include <vector>

namespace scopi_hip {

// Layouts of the contact Jacobian B (one block row per contact, two particle blocks).
//   kRotFree : NoFriction with spherical active particles.  The contact normal passes
//              through the sphere centre, so the rotational blocks R^T (r x n) vanish:
//              row c = [ n ] on particle i, -[ n ] on j (3 values per contact).
//   kRigid   : NoFriction, general shapes: row c = [ n | a_i ] on i, -[ n | a_j ] on j,
//              a_k = R_k^T (r_k x n), r_k = p_k - x_k (9 values per contact).
//   kCone    : Friction: rows (c, d), d = 0..2, = [ e_d | K_i[d] ] on i, -[ e_d | K_j[d] ] on j,
//              K_k[d] = R_k^T (r_k x e_d); n is kept for the cone projection (21 values).
//   kVisc    : Viscous / ViscousFriction: B as kCone, and 4 multiplier slots per contact that
//              hold the contact's 1, 2 or 4 rows of the reference's local vector (unused slots
//              stay 0).  The per-contact kind says how they map to the 3 global rows:
//                kind 0: lambda = x0 n (x0 >= 0);
//                kind 1: lambda = (x0 - x1) n (x0, x1 >= 0);
//                kind 2: lambda = -x0 n + (x1, x2, x3) (the nearer of x0 >= 0 and the cone).
enum Layout : int { kRotFree = 0, kRigid = 1, kCone = 2, kVisc = 3 };

template <int LAY>
struct lay;
template <>
struct lay<kRotFree> {
    static constexpr int NR = 1, NC = 3, NT = 3, NW = 3;
};
template <>
struct lay<kRigid> {
    static constexpr int NR = 1, NC = 9, NT = 6, NW = 6;
};
template <>
struct lay<kCone> {
    static constexpr int NR = 3, NC = 21, NT = 9, NW = 6;
};
template <>
struct lay<kVisc> {
    static constexpr int NR = 4, NC = 21, NT = 9, NW = 6;
};
// kVisc kind word: the kind in bits 0-1, bit 2 set when gamma == gamma_min
constexpr int kKindMask = 3, kAtGammaMin = 4;

// Scalar state of the projected-gradient loop, owned by the device.  Every block of
// the next round's trial kernel runs the reference's control logic
// (include/scopi/solvers/apgd.hpp:164-213) on its own copy after each product, so
// the host only polls `done`.  Two copies alternate: a round reads one and its
// trial kernel's block 0 publishes the updated state in the other.
struct ApgdCtrl {
    double L, alpha, theta, beta;
    double L_acc;        // L of the last accepted trial (before the 0.97 decay)
    double tol, rtol, dt2;
    double res;          // last ||lambda_{k+1} - lambda_k||
    double best_res;     // FP32 phase: smallest res so far, to within a factor 2
    int64_t since_best;  // FP32 phase: accepted iterations since best_res last halved
    double red[9];       // sums of the trial kernel (0..6, 8) and <lambda+, Q lambda+> from k_bt (7)
    int64_t iter, backtracks, max_ite, q_applies, consec_bt;
    int32_t cur, prev, trial;
    int32_t done, status, method, dynamic, low_precision;
    int32_t pending;     // a product waits for its control step (set by k_bt)
};

// Independent scenes solved together.  The scenes' active particles and contacts
// occupy consecutive ranges of one packed problem; the per-contact kernel
// (k_trial) and k_bt run one grid over all scenes in which every block
// belongs to one scene.  Each scene keeps its own control state (ApgdCtrl) and
// block sums, and its blocks split its rows the same way whichever other scenes
// share the launch, so a scene's iterates do not depend on the others.
// A null `rblk` means one scene spanning the whole grid.
struct SceneMap {
    const int32_t* rblk = nullptr;     // per-contact grid: scene of each block
    const int32_t* rfirst = nullptr;   // per scene: its first block in the per-contact grid
    const int32_t* rcount = nullptr;   // per scene: its blocks in the per-contact grid
    const int32_t* bblk = nullptr;     // k_bt grid: scene of each block
    const int32_t* bfirst = nullptr;   // per scene: its first block in the k_bt grid
    const int32_t* bcount = nullptr;   // per scene: its blocks in the k_bt grid
    const int32_t* slist = nullptr;    // scenes of the grid, in order (one control block each)
    const int64_t* crange = nullptr;   // per scene: contacts [lo, hi)
    const int64_t* arange = nullptr;   // per scene: active particles [lo, hi)
    int32_t* done = nullptr;           // per scene: done flag, written by the control steps
    int64_t pstride = 0;               // trial kernel's block sums: quantity k of block b at k * pstride + b
                                       // (0: kRedBlocksMax, the one-scene buffer)
};

// Final state of one scene solved by the per-scene work-queue kernel (k_apgd_queue).
struct QueueResult {
    int64_t iter, iter32, backtracks, q_applies;
    double res, L_acc;
    int32_t status, pad;
};

// Arguments of the work-queue kernel: the scenes it solves, largest first, and the solve's
// parameters (the FP32 phase's stop for FP32 + FP64 refinement).
struct QueueArgs {
    const int32_t* order = nullptr;   // scenes to solve, in the order they are taken
    int32_t nqueue = 0;
    int32_t* next = nullptr;          // work counter, 0 at launch
    const int64_t* crange = nullptr;  // per scene: contacts [lo, hi)
    const int64_t* arange = nullptr;  // per scene: active particles [lo, hi)
    QueueResult* result = nullptr;    // per scene
    const int* tstart = nullptr;      // particle-major copy of B: entries of each active particle
    const int32_t* tc = nullptr;
    const int32_t* ci = nullptr;
    const int32_t* cj = nullptr;
    int64_t ni = 0;                   // obstacles before the active particles
    double alpha = 0, tol = 0, tol32 = 0, rtol32 = 0, dt2 = 0;
    int64_t max_ite = 0;
    int32_t method = 0, dynamic = 0, warm = 0, fp32 = 0;
};

// Host description of a scene set (scenes in order, consecutive ranges).
struct SceneSet {
    int nscene = 0;
    const int64_t* d_crange = nullptr;   // device, 2 per scene: contacts [lo, hi)
    std::vector<int64_t> alo, ahi;       // active particles of each scene (active index)
};

// Device arrays of one precision.  Per-contact and per-particle data are
// arrays of small structures (AoS): coef[c*NC + k], vectors[c*NR + d], w[p*NW + k].
template <class T>
struct BMatrix {
    int64_t cap = 0;     // contacts
    int64_t pcap = 0;    // particles
    T* coef = nullptr;   // row-major coefficients, NC per contact
    T* tcoef = nullptr;  // particle-major copy, NT per entry (2 entries per contact), sign folded in
    T* c = nullptr;      // g2l(D n + dt A U), NR per contact
    T* mu = nullptr;     // friction coefficient per contact
    T* minv = nullptr;   // diagonal of M^-1, NW per particle
    T* lam = nullptr;    // 3 slots of NR*cap
    T* qlam = nullptr;   // 3 slots of NR*cap: Q lambda_s, written by the rounds that read lambda_s as lambda_k
    T* w = nullptr;      // 3 slots of 6*pcap: w_s = M^-1 B^T lambda_s, NW per particle
    const int32_t* kind = nullptr;  // kVisc: kind word per contact (shared by both precisions)
};

class Solver {
  public:
    ~Solver();
    // Assembles B and c from the device contacts and the a-priori velocities,
    // solves, and leaves u = v + dt M^-1 B^T lambda in u() / w() (3 x pcap each).
    // rot_free: every active particle is a sphere (enables the kRotFree layout).
    // Viscous / ViscousFriction: `gamma` (device, nc) is each contact's adhesion potential; the
    // solve reads it for the multiplier layout and leaves the updated values in it.
    // lam0 (device, NR per contact, FP64; may be null): the first iterate instead of lambda = 0 (NoFriction
    // and Friction).
    int solve(const Particles& P, const Contacts& C, int64_t nc, const scopi_hip_params& prm, double dt,
              bool rot_free, hipStream_t s, scopi_hip_solve_info* info, double* gamma = nullptr,
              const double* lam0 = nullptr);
    // the last solve's multipliers (device, FP64, rows per contact rows_per_contact())
    const double* final_lambda() const { return b64_.lam + (int64_t)final_slot_ * nr_ * b64_.cap; }
    int rows_per_contact() const { return nr_; }
    // The same for a set of independent scenes packed in P / C (scene ranges in `S`);
    // one solve info per scene.  Afterwards every scene's final lambda is in slot 0.
    // lam0 (NR rows per contact, FP64): every scene's first iterate (warm start), or null for 0.
    int solve_scenes(const Particles& P, const Contacts& C, int64_t nc, const scopi_hip_params& prm, double dt,
                     bool rot_free, hipStream_t s, const SceneSet& S, std::vector<scopi_hip_solve_info>& info,
                     const double* lam0 = nullptr);
    // lambda of the last solve (FP64) in the reference's order (last_rows() values).
    void download_lambda(double* host, hipStream_t s) const;
    // rows [NR lo, NR hi) of the last solve's lambda (FP64)
    void download_lambda_range(double* host, int64_t lo, int64_t hi, hipStream_t s) const;
    const double* u() const { return u_; }
    const double* w() const { return w_; }
    int64_t pcap() const { return pcap_; }
    int64_t last_rows() const { return rows_; }
    bool has_solution() const { return has_solution_; }

    void set_timing(bool on) {
        timing_ = on;
        t_bt_ms_ = t_row_ms_ = t_trial_ms_ = 0.0;
        n_bt_ = n_row_ = n_trial_ = 0;
    }
    // [ms_bt, n_bt, ms_row, n_row, ms_trial, n_trial]; the row product runs inside k_trial,
    // so ms_row and n_row stay 0
    void timing(double* out) const;
    int layout() const { return layout_; }
    // launch-shape tuning (SCOPI_TUNE_* keys of scopi_hip.h)
    int set_tuning(int key, int value);

  private:
    void reserve(int64_t na, int64_t nc, bool fp32);
    template <class T, int LAY>
    void run_phase(BMatrix<T>& B, const Particles& P, const Contacts& C, int64_t na, int64_t nc,
                   const scopi_hip_params& prm, double dt, double tol, double rtol, int64_t max_ite, bool warm,
                   bool keep_L, hipStream_t s, scopi_hip_solve_info& info);
    template <int LAY>
    void solve_layout(const Particles& P, const Contacts& C, int64_t nc, const scopi_hip_params& prm, double dt,
                      hipStream_t s, scopi_hip_solve_info& info);
    template <int LAY>
    void assemble(const Particles& P, const Contacts& C, int64_t nc, double dt, bool fp32, hipStream_t s);
    template <int LAY>
    void solve_phases(const Particles& P, const Contacts& C, int64_t nc, const scopi_hip_params& prm, double dt,
                      hipStream_t s, scopi_hip_solve_info& info);
    template <int LAY>
    void solve_fixed_point(const Particles& P, const Contacts& C, int64_t nc, const scopi_hip_params& prm, double dt,
                           hipStream_t s, scopi_hip_solve_info& info);
    // Viscous problems: kind words of the contacts and the multiplier rows they make
    int64_t viscous_kinds(const double* gamma, int64_t nc, const scopi_hip_params& prm, hipStream_t s);
    // multi-scene loop: tables of the scenes still running, phase loop, per-scene results
    void scene_tables(const std::vector<int>& scenes, hipStream_t s);
    template <class T, int LAY>
    void run_phase_scenes(BMatrix<T>& B, const Contacts& C, int64_t na, int64_t nc, const scopi_hip_params& prm,
                          double dt, double tol, double rtol, int64_t max_ite, bool warm, bool keep_L, hipStream_t s,
                          const std::vector<int>& scenes, std::vector<scopi_hip_solve_info>& info);
    // the scenes of `scenes` solved whole by the per-scene work-queue kernel (both phases), on the
    // queue stream after the work enqueued so far on `s`; queue_join makes `s` wait for it and
    // reads back the scenes' results
    template <int LAY>
    void queue_launch(const Contacts& C, const scopi_hip_params& prm, double dt, hipStream_t s,
                      const std::vector<int>& scenes);
    void queue_join(hipStream_t s, const std::vector<int>& scenes, std::vector<scopi_hip_solve_info>& info);
    template <int LAY>
    void solve_layout_scenes(const Particles& P, const Contacts& C, int64_t nc, const scopi_hip_params& prm, double dt,
                             hipStream_t s, std::vector<scopi_hip_solve_info>& info);

    RadixSorter sorter_;
    Scanner scanner_;
    int64_t pcap_ = 0, ccap_ = 0;
    BMatrix<double> b64_;
    BMatrix<float> b32_;
    bool b32_alloc_ = false;
    // particle-major transpose: contacts where the particle is i (contiguous, sorted
    // by (i, j)) and where it is j (stable radix sort), merged into one entry list
    int* istart_ = nullptr;
    int* iend_ = nullptr;
    int* jstart_ = nullptr;
    int* jend_ = nullptr;
    int* tcount_ = nullptr;
    int* tstart_ = nullptr;   // na + 1
    int32_t* tc_ = nullptr;   // 2 cap: contact index, ~c for the j side
    uint32_t* jkeys_ = nullptr;
    int32_t* jlist_ = nullptr;
    double* u_ = nullptr;  // 3 x pcap
    double* w_ = nullptr;  // 3 x pcap
    ApgdCtrl* ctrl_ = nullptr;    // 2 alternating copies
    ApgdCtrl* h_ctrl_ = nullptr;  // pinned
    double* partial_trial_ = nullptr;  // 8 block sums per block of k_trial, quantity-major (SceneMap::pstride)
    double* partial_bt_ = nullptr;     // one per block of k_bt, contiguous
    int64_t partial_bt_cap_ = 0;
    int64_t rows_ = 0, nc_ = 0;
    int nr_ = 1;
    int final_slot_ = 0;
    const double* lam0_ = nullptr;  // first iterate of the current solve (warm start), or null
    int ctrl_cur_ = 0;  // copy of ApgdCtrl with the last phase's final state
    int layout_ = kRotFree;
    int last_problem_ = SCOPI_PROBLEM_NOFRICTION;
    int reduce_blocks_ = 1024;  // grid of the per-contact kernels (grid-stride)
    int bt_group_ = 8;          // lanes per particle in k_bt
    int row_iters_ = 2;         // loop iterations per lane of the per-contact grid below its cap
    int graph_rounds_ = 0;      // > 0: the one-scene loop replays a hipGraph of this many rounds
    hipGraphExec_t graph_exec32_ = nullptr, graph_exec64_ = nullptr;  // kept across solves (updated in place)
    int64_t last_rounds_[2] = {0, 0};  // rounds of the last one-scene solve phase, FP64 [0] and FP32 [1]
    bool has_solution_ = false;
    // FrictionFixedPoint: c without S, s_ij, block sums of the error
    double* fp_c0_ = nullptr;
    double* fp_s_ = nullptr;
    double* fp_partial_ = nullptr;
    double* h_fp_partial_ = nullptr;
    int64_t fp_cap_ = 0;
    // Viscous problems: kind words, the row count of a solve (device + pinned)
    int32_t* vkind_ = nullptr;
    int64_t vkind_cap_ = 0;
    unsigned long long* vrows_ = nullptr;
    unsigned long long* h_vrows_ = nullptr;
    // multi-scene state
    const SceneSet* scenes_ = nullptr;
    int nscene_cap_ = 0;
    int64_t* d_arange_ = nullptr;    // 2 per scene
    int32_t* d_done_ = nullptr;      // per scene
    int32_t* h_done_ = nullptr;      // pinned, per scene
    int32_t* d_buf_ = nullptr;       // per scene: ctrl copy holding its state (phase end)
    int32_t* d_tab_ = nullptr;       // block tables (SceneMap arrays)
    int32_t* h_tab_ = nullptr;       // pinned staging of the tables
    int64_t tab_cap_ = 0;
    ApgdCtrl* h_ctrls_ = nullptr;    // pinned, per scene
    ApgdCtrl* mctrl_ = nullptr;      // 2 copies x scenes
    int64_t partial_cap_ = 0;        // blocks
    int64_t p_inactive_ = 0;
    SceneMap map_;
    int rgrid_ = 0, bgrid_ = 0, nlist_ = 0;
    std::vector<int> scene_rblocks_, scene_bblocks_;
    std::vector<int32_t> scene_buf_;  // per scene: ctrl copy holding its final state
    std::vector<int64_t> scene_nc_;   // per scene: contacts of the current solve
    // per-scene work queue: scenes of at most queue_max_contacts_ contacts (0: none) are solved by
    // k_apgd_queue, queue_blocks_ workgroups of queue_threads_ threads at most
    int64_t queue_max_contacts_ = 0;
    int queue_blocks_ = 1024;
    int queue_threads_ = 256;
    hipStream_t qstream_ = nullptr;   // the work-queue launches (default priority)
    hipEvent_t qfork_ = nullptr, qjoin_ = nullptr;
    int32_t* d_qorder_ = nullptr;
    int32_t* h_qorder_ = nullptr;     // pinned
    int32_t* d_qnext_ = nullptr;
    QueueResult* d_qres_ = nullptr;
    QueueResult* h_qres_ = nullptr;   // pinned
    int qcap_ = 0;
    // kernel timing (HIP events on the solve stream)
    bool timing_ = false;
    double t_bt_ms_ = 0, t_row_ms_ = 0, t_trial_ms_ = 0;
    int64_t n_bt_ = 0, n_row_ = 0, n_trial_ = 0;
};

}  // namespace scopi_hip
