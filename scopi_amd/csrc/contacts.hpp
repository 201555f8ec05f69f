// contacts.hpp -- host side of the device contact search.
#pragma once

#include "common.hpp"
#include "sort_scan.hpp"
#include "superellipsoid.hpp"

#include <vector>

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
    // bounded: every searched particle has a finite bounding radius (false when periodic copies
    // of planes are searched as particles: the candidate cut-off is then the kd-tree radius, or
    // every pair for brute force).
    int64_t run(const Particles& P, Contacts& C, const scopi_hip_params& prm, hipStream_t s, bool closed_form = true,
                bool bounded = true);
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
    int32_t* keep_ = nullptr;  // the count pass's first partners of each particle (kKeepPartners each)
    int64_t pair_cap_ = 0;
    int* d_totals_ = nullptr;   // [0] active pairs, [1] obstacle pairs, [2] a particle had more partners than it kept
    int* h_totals_ = nullptr;   // pinned
    int64_t nc_obs_ = 0;
    SeGrids se_grids_;
    SeCandidates se_cand_;
    int* cand_flags_ = nullptr;
    int* cand_offsets_ = nullptr;
    int64_t cand_cap_ = 0, ncand_ = 0;
};

// Ranges of one scene of a batch.  A batch packs the obstacles of all scenes
// first, then the active particles of all scenes, each scene in a consecutive
// range of both; its contacts come out in one consecutive range, obstacle pairs
// first, each part in the reference's (i, j) order.
struct SceneRanges {
    int64_t ob_lo, ob_hi;   // obstacles (particle indices)
    int64_t a_lo, a_hi;     // active particles (particle indices)
    int64_t u_lo;           // first work unit: obstacle x active units, then one unit per active particle
    int64_t cell0, cells;   // the scene's range of cell keys
};

// Contact search of every scene of a batch in one pass: per-scene uniform
// grids (dense cell keys in the scene's own key range, at most `cells` cells),
// one radix sort over all active particles, one count / scan / emit over the
// work units of all scenes.  Each scene's pair set is exactly the one
// ContactSearch::run finds for that scene alone (closed-form shapes).
class BatchSearch {
  public:
    ~BatchSearch();
    // Uploads the scene table (host order = packed order); call after repacking.
    void set_scenes(const std::vector<SceneRanges>& scenes, int64_t n_inactive, int64_t na, hipStream_t s);
    // Returns nc (host syncs inside); per-scene contact ranges in d_crange() / crange().
    // closed_form false: superellipsoids present, candidate path with the device closest points.
    int64_t run(const Particles& P, Contacts& C, const scopi_hip_params& prm, hipStream_t s, bool closed_form = true);
    const int64_t* d_crange() const { return d_crange_; }
    const std::vector<int64_t>& crange() const { return h_crange_; }

  private:
    RadixSorter sorter_;
    Scanner scanner_;
    int K_ = 0;
    int64_t na_ = 0, ni_ = 0, units_ = 0, cells_ = 0;
    SceneRanges* d_scenes_ = nullptr;
    int64_t* d_ulo_ = nullptr;       // K + 1 unit offsets
    int32_t* d_scene_of_ = nullptr;  // scene of each active particle
    GridInfo* d_grids_ = nullptr;
    int64_t* d_crange_ = nullptr;    // 2 K
    int64_t* h_io_ = nullptr;        // pinned: total + 2 K
    std::vector<int64_t> h_crange_;
    int kcap_ = 0;
    uint32_t* keys_ = nullptr;
    int32_t* vals_ = nullptr;
    int64_t pcap_ = 0;
    int* counts_ = nullptr;
    int* offsets_ = nullptr;
    int64_t ucap_ = 0;
    int* cell_start_ = nullptr;
    int* cell_end_ = nullptr;
    int64_t ccap_ = 0;
    int* d_total_ = nullptr;
    SeGrids se_grids_;
    SeCandidates se_cand_;
    int* cand_flags_ = nullptr;
    int* cand_offsets_ = nullptr;
    int64_t cand_cap_ = 0;
};

// Allocates / grows the structure-of-arrays contact list (keeps nothing).
void contacts_reserve(Contacts& C, int64_t n);
void contacts_free(Contacts& C);

}  // namespace scopi_hip
