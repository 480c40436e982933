// context.hpp -- the engine context and host helpers shared by engine.hip
// (plain loop) and learn.hip (guided loop).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "common.hpp"

namespace cuhre {

struct BatchEngine;
void guide_detach(cuhre_guide* g);  // the guide's context is going away

struct SplitArgs {
  const double* lower;
  const double* width;
  int64_t ld;
  int64_t n;
  const double* est;
  const double* err;
  const int8_t* axis;
  double epsrel, epsabs;
  const int64_t* offsets;  // per evaluation block (rpb regions)
  int32_t rpb;
  int32_t dim;
  double* out_lower;
  double* out_width;
  int64_t out_ld;
  const IterState* state;  // nullptr: always split
  const int8_t* retire;    // optional: regions retired as finished are not split
  const int* retire_on;    // optional device flag: retire honoured only when nonzero
  const int64_t* gbase;    // optional: base offset per group of 64 blocks (two-level finalisation)
};

void launch_split(const SplitArgs& S, hipStream_t st);

}  // namespace cuhre

struct cuhre_context {
  int device = 0;
  hipStream_t stream = nullptr;
  int64_t capacity = 0;        // regions per list; 0: three quarters of free HBM
  int64_t batch_capacity = 0;  // 0: sized from free HBM per batch call
  int profile = 0;
  int segments = 0;                // 0: automatic, else 1, 2, 4, 8 or 16
  int64_t fill_lanes = 262144;     // auto: smallest S with n*S >= fill_lanes
  int64_t half_max_regions = 8192;  // auto: half-wavefront form up to this many regions

  int64_t cap = 0;  // allocated regions per list buffer
  // free HBM plus what the list buffers hold (bytes), measured when first
  // needed and again after the lists are reallocated (< 0: measure)
  double avail_bytes = -1.0;
  int dim_alloc = 0;
  double* lo[2] = {nullptr, nullptr};
  double* wd[2] = {nullptr, nullptr};
  double* est = nullptr;
  double* err = nullptr;
  double* abssum = nullptr;
  int8_t* axis = nullptr;
  cuhre::Partial* partials = nullptr;
  int64_t* offsets = nullptr;
  cuhre::IterState* state = nullptr;
  cuhre::IterState* h_state = nullptr;
  // coherent host memory the finalisation kernel publishes the state to
  // (host and device views; two buffers, the launch numbered seq uses
  // buffer seq & 1), with the sequence number written last
  cuhre::IterState* pub_state = nullptr;
  uint32_t* pub_seq = nullptr;
  cuhre::IterState* pub_state_dev = nullptr;
  uint32_t* pub_seq_dev = nullptr;
  uint32_t seq = 0;
  // plain loop: queue each iteration before the previous state is read,
  // with the list lengths and finished sums in these device slots (one per
  // list buffer)
  int queue_ahead = 1;
  cuhre::SpecSlot* spec = nullptr;
  // half-wavefront iterations queued ahead alternate between these two sets
  // of block partials, dense counts and active-region lists, so an
  // evaluation can split its parent list while it writes its own
  struct FuseSet {
    cuhre::Partial* part = nullptr;  // kOneLevel blocks
    int64_t* cnt = nullptr;
    uint8_t* act = nullptr;  // kOneLevel * 8 regions
  } fuse[2];
  char* fuse_mem = nullptr;
  double* aux = nullptr;  // i5 (cap) + dvec (cap*dim) on request
  int64_t aux_cap = 0;
  char* p2_ws = nullptr;  // Phase-2 workspace of the two-phase method, kept between calls
  size_t p2_bytes = 0;
  double2* sum_part = nullptr;
  double* sum_out = nullptr;

  double* table = nullptr;
  std::vector<double> table_host;
  float* ftable = nullptr;  // LC_LT_t nodes and fields (float, as the reference's textures)
  std::vector<float> ftable_host;

  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  cuhre::BatchEngine* batch = nullptr;
  std::vector<cuhre_guide*> guides;  // attached guides, detached on destroy
};

namespace cuhre {

// Error reporting: sets the thread-local message cuhre_last_error returns
// and returns code.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail_msg(int code, const char* msg);
int check_integrand_args(const cuhre_integrand* f, int dim, int key);
int check_bounds(int dim, double epsrel, double epsabs, const double* lows, const double* highs);
double region_flops(const cuhre_integrand* f, int d, int key);
double family_flops(const cuhre_integrand* f, int d);
double ipow_flops(int e, int d);
int grid_k(int d);  // initial grid k^d: k = 4 (d <= 4), 2 (5..10), 1 above (GPUKernelquad.cu:300-316)

// The static workspace of cuhre_integrate_noalloc: 64 MiB per device, the
// last 1 MiB for integrand tables.
constexpr size_t kScratchBytes = 64ull << 20;
constexpr size_t kScratchTableBytes = 1ull << 20;
size_t al256(size_t b);
size_t scratch_list_bytes(int64_t cap, int dim);  // list buffers of cap regions
int64_t scratch_capacity(int dim);                // regions per list the workspace holds
int ctx_reserve(cuhre_context* c, int64_t cap, int dim, int64_t keep_n = 0, int cur = 0, int64_t keep_nb = 0);
int ctx_reserve_aux(cuhre_context* c, int64_t cap, int dim);
// Device copy of an integrand.  Tabulated families copy their table to the
// device: into buffers the context allocates and keeps (table_dst null), or
// into table_dst, a caller-provided device area of table_bytes (the
// no-allocation entry point), in which case nothing is allocated.
int make_dev_integrand(cuhre_context* c, const cuhre_integrand* f, int dim, DevIntegrand* P,
                       char* table_dst = nullptr, size_t table_bytes = 0);
int64_t list_capacity(cuhre_context* c, int dim);
int pick_segments(const cuhre_context* c, int64_t n, int key = 7, int family = -1, int dim = 0);
// Block partials / offsets the list buffers of cap regions come with: one
// per 64 regions, and at least 1024 (the half-wavefront form on 8192 regions).
inline int64_t block_capacity(int64_t cap) { return cap / kWave + 1 > 1024 ? cap / kWave + 1 : 1024; }
// Group sums of the two-level finalisation: one per 64 blocks.
inline int64_t group_capacity(int64_t nbmax) { return nbmax / kWave + 2; }
// Regions of the initial k^d grid, and the grid written into list `list`
// on the device.
int64_t initial_grid_count(int d);
int upload_initial_grid(cuhre_context* c, int list, int d, const double* lows, const double* highs, int64_t n0);
int check_common(cuhre_context* ctx, const cuhre_integrand* f, int dim, int key);
void fill_eval_args(cuhre_context* c, EvalArgs& A, int cur, int64_t n, int dim, double epsrel, double epsabs);
void fill_split_args(cuhre_context* c, SplitArgs& S, int cur, int64_t n, int dim, double epsrel, double epsabs,
                     int rpb, bool use_state);
void dd_add_host(double& s1, double& c1, double s2, double c2);
int64_t blocks_for(int64_t n, int S);
RuleConst make_rule(int d, int key = 7);
int gm_points(int d);
int r9_points(int d);
int set_error(int code, const char* msg);

}  // namespace cuhre
