// batch.h -- many independent gel-fluid instances stepped as one.
//
// Extension with no counterpart in the reference (whose drivers step one
// instance: sim.cpp:60-111).  A Batch holds instances made of the
// reference's own objects -- a gel vector, its Coupler and its Fluid -- and
// advances all of them by one time step with one launch per phase: the
// gel phases over every gel, the marker phases over every marker, and the
// 25 stream/collide sub-steps as one persistent launch each over the
// concatenated 256-cell blocks of every lattice.  Each instance gives the
// same bits as stepping its objects alone (sim.cpp's loop body); only the
// launch shape differs.
//
// The step can be captured in a HIP graph (one per buffer parity) and the
// stream/collide launch shape chosen by timing the candidates.
#pragma once

#include <map>
#include <memory>
#include <vector>

#include "coupling.h"
#include "fluid.h"
#include "gel.h"
#include "output.h"

namespace glsm {
struct BatchDevice;
}

class Batch {
 public:
  // storage: the distribution storage of every member (Fluid::Storage)
  explicit Batch(int storage = Fluid::F32_DEV, int scale_exp = 20);
  ~Batch();
  Batch(const Batch&) = delete;
  Batch& operator=(const Batch&) = delete;

  // Adds an instance.  The objects stay owned by the caller and must outlive
  // its membership; `gels` must be the vector the coupler was built from.
  // The fluid is converted to the batch's storage.  `iter` is the iteration
  // of the instance's next step.  With write_files the .dat output of
  // Gel::writeFiles / Fluid::writeFiles is produced at the reference's
  // cadence after each step.  Returns the member id.
  // With an output stream attached (setOutput) the files go through it,
  // into `out_dir`; without one, through the objects' own writeFiles.
  int add(std::vector<Gel*>& gels, Coupler* coupler, Fluid* fluid, int iter, bool write_files = false,
          const std::string& out_dir = "");
  void remove(int id);
  int size() const { return (int)members_.size(); }
  std::vector<int> ids() const;
  int iteration(int id) const;  // the member's next iteration
  long cells() const;           // lattice cells over all members

  // `nsteps` time steps of every member (sim.cpp:66-81 each).
  void step(int nsteps = 1);

  // Capture each step in a HIP graph (default on).
  void useGraph(bool on);

  // Asynchronous .dat output for members added with write_files (nullptr =
  // the objects' synchronous writeFiles).  The stream must outlive its use.
  void setOutput(OutputStream* out) { output_ = out; }

  // Launch shape of the batched stream/collide kernel.
  struct Launch {
    int nt;         // non-temporal stores
    int run;        // XCD run length: 0 (one contiguous run per XCD), 4, 16 or 64
    int grid;       // work-groups; 0 = one per work unit, else persistent
    int wgb = 1;    // 256-cell blocks stepped at once by a work-group: 1, 2, 4 (256-1024 threads)
    int chunk = 1;  // rounds of blocks per work-group (cells per thread), 1..64
    int lds = 0;    // LDS bytes reserved per work-group (0..65536; one block per work-group only):
                    // caps the work-groups resident per CU at 160 KB / lds
  };
  Launch launch() const;
  void setLaunch(const Launch& l);
  // Times every candidate shape on the members' current state (state is
  // not changed; each shape's time is the median of three trials of `reps`
  // launches) and keeps the fastest; returns it and its mean launch time.
  // Candidates: NT stores x XCD run length x grid (one work-group per unit,
  // or 2-16 per CU persistent) with one block per work-group, the
  // one-work-group-per-block shapes also with 32, 40 and 52 KB of LDS
  // reserved (5, 4 and 3 work-groups per CU), and work-group sizes of 512 /
  // 1024 threads and 2-8 rounds of blocks per work-group (one work-group
  // per unit).
  Launch autotune(int reps, double* best_ms);

  // Mean duration of the batched stream/collide kernel (start/stop events
  // timestamped by the dispatch, middle sub-step of every `every`-th step;
  // 0 = off).
  void enableLbmTiming(int every);
  double lbmKernelTiming(int* samples);

 private:
  struct Member {
    int id;
    std::vector<Gel*>* gels;
    Coupler* coupler;
    Fluid* fluid;
    int iter;
    bool write_files;
    std::string out_dir;
    // the fluid buffers bound into the device tables, by parity: a member
    // stepped outside the batch, or given new buffers, no longer matches and
    // the tables are rebuilt
    void* ftab[2];
    float* ctab[2];
    // the member's parameters as the device tables hold them (fluid, IBM,
    // coupler and gel values); a change made through the objects' own
    // setters between steps is seen and the tables are rebuilt
    std::vector<char> params;
  };
  std::vector<char> memberParams(const Member& m) const;
  void rebuild();
  bool listsNeeded() const;
  void enqueueStep(bool timed, bool lists);
  void dropGraphs();

  int storage_;
  int scale_exp_;
  int next_id_ = 1;
  std::vector<Member> members_;
  bool dirty_ = true;
  bool use_graph_ = true;
  bool launch_set_ = false;
  int pf_ = 0;  // which of a member's two distribution buffers holds its state
  int pc_ = 0;  // which of its two concentration buffers holds its state
  int nsub_ = 0;
  Launch launch_{0, 0, 0, 1, 1};
  std::unique_ptr<glsm::BatchDevice> dev_;
  bool lists_fresh_ = false;        // IBM lists match the members' marker positions
  std::vector<uint64_t> seen_pos_;  // gels' position versions after the last step
  std::map<int, void*> graphs_;     // (pf, pc, lists) -> hipGraphExec_t
  int timing_every_ = 0;
  long timing_calls_ = 0;
  std::vector<void*> timing_events_;
  OutputStream* output_ = nullptr;
};
