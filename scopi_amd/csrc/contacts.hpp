// contacts.hpp -- host side of the device contact search.
#pragma once

#include "common.hpp"
#include "sort_scan.hpp"
#include "superellipsoid.hpp"

namespace scopi_hip {

struct GridInfo {
    double lo[3];
    double w[3];     // cell widths (>= the candidate cut-off)
    int nc[3];       // cells per axis
    int bits;        // Morton bits per axis
    double cutoff;   // candidate cut-off (centre distance)
    double rmax;     // largest bounding radius among active particles
};

// Uniform-grid candidate search that emits exactly the pair set of
// contact_kdtree::run_impl / contact_brute_force::run_impl.
class ContactSearch {
  public:
    ~ContactSearch();
    // Runs on `s`; grows `C` when needed; returns nc (host sync inside).
    // closed_form: every shape pair has a closed-form distance (no superellipsoid).
    int64_t run(const Particles& P, Contacts& C, const scopi_hip_params& prm, hipStream_t s, bool closed_form = true);
    int64_t last_obstacle_contacts() const { return nc_obs_; }
    // superellipsoid path: candidates of the last run and their closest points
    int64_t last_candidates() const { return ncand_; }
    const SeCandidates& candidates() const { return se_cand_; }

  private:
    int64_t run_candidates(const Particles& P, Contacts& C, const scopi_hip_params& prm, hipStream_t s);
    void reserve_particles(int64_t na, int64_t n_inactive);
    void reserve_pairs(int64_t n);
    RadixSorter sorter_;
    Scanner scanner_;
    GridInfo* d_grid_ = nullptr;
    double* d_partial_ = nullptr;
    int64_t partial_cap_ = 0;
    uint32_t* keys_ = nullptr;
    int32_t* vals_ = nullptr;
    int* counts_ = nullptr;
    int* offsets_ = nullptr;
    int* obst_flags_ = nullptr;
    int* obst_offsets_ = nullptr;
    int64_t pcap_ = 0, ocap_ = 0;
    int* cell_start_ = nullptr;
    int* cell_end_ = nullptr;
    int64_t cells_cap_ = 0;
    int32_t* pair_j_ = nullptr;
    int64_t pair_cap_ = 0;
    int* d_totals_ = nullptr;   // [0] active pairs, [1] obstacle pairs
    int* h_totals_ = nullptr;   // pinned
    int64_t nc_obs_ = 0;
    SeGrids se_grids_;
    SeCandidates se_cand_;
    int* cand_flags_ = nullptr;
    int* cand_offsets_ = nullptr;
    int64_t cand_cap_ = 0, ncand_ = 0;
};

// Allocates / grows the structure-of-arrays contact list (keeps nothing).
void contacts_reserve(Contacts& C, int64_t n);
void contacts_free(Contacts& C);

}  // namespace scopi_hip
