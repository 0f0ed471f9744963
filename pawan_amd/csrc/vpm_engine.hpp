// vpm_engine.hpp -- host side of the gfx950 interaction path: device-resident
// jobs, the work-item plan, kernel dispatch, HIP-graph replay of a step, and
// timing.  Shared by the single-simulation context, the batch and the server.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "vpm_types.hpp"

namespace pw {

// Error carrying a PW_E* status; converted to a return code at the C ABI.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char *what);
#define PW_HIP(x) ::pw::hip_check((x), #x)

// Owned device allocation.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t b);
    ~DevBuf();
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    void reset(size_t b);  // reallocate (contents lost)
    template <typename T>
    T *as() const { return (T *)p; }
};

// Pinned host staging buffer.
struct HostBuf {
    void *p = nullptr;
    size_t bytes = 0;
    HostBuf() = default;
    ~HostBuf();
    void ensure(size_t b);
    template <typename T>
    T *as() const { return (T *)p; }
};

struct JobParams {
    double nu = 2.5e-3;
    double dt = 0.0;
    double uinf[3] = {0, 0, 0};
    int precision = PREC_FP64;
    int chunks = 0;  // source chunks of the job (0: the chunk policy); see chunk_len
};

// Source chunk length for a job of n particles with TB targets per tile:
// 16 chunks per job, or up to one tile per chunk when the tiles number fewer
// than 256 (about 4,096 tile x chunk items); chunk a multiple of TB.  Depends
// on (n, TB) only, so a job's sums are formed the same way alone or in any
// batch.
size_t chunk_for(size_t n, int TB);
// Source chunk length for a job on the symmetric pass (blocks of TB x TPT
// targets): a whole number of blocks, so that the chunks and the strips
// (strip_for_sym) are the same 128 (or, past 1,024 blocks, blocks / 8)
// stretches of particles.  The upper triangle of (strip, chunk) work items
// then numbers ~8,256 or more (16 per workgroup of a full persistent grid),
// and a particle's partial sums take about as many target-side rows as
// source-side ones (~128 in all).  Depends on (n, TB, TPT) only.
size_t chunk_for_sym(size_t n, int TB, int TPT);
// Source chunk length of a job of n particles.  chunks > 0 asks for about that
// many chunks: the length is rounded up to whole blocks of TB x TPT targets on
// the symmetric pass (chunks and strips coincide there), to whole tiles of TB
// one-sided, so the count is at most `chunks`.  chunks == 0: the policy above.
// Depends on its arguments only.
size_t chunk_len(size_t n, int TB, int TPT, bool sym, int chunks);
// Strip width of a symmetric job with source chunks of L (a whole number of
// blocks of TB x TPT targets, about L): a work item is a strip's blocks
// against one chunk, and the strip's source-side sums fill one slot row.
size_t strip_for_sym(size_t L, int TB, int TPT);

// Bytes a resident job occupies on the device (TB x TPT targets per block;
// sym: the job uses the symmetric pass and its source-side slots).
size_t job_footprint(size_t n, int precision, int TB, int TPT, bool sym, int chunks = 0);
// Threads per workgroup of a precision's jobs in an engine built with `block`:
// 512 exists only for the binary32 symmetric HIGHORDER pass at one target per
// thread (half the source-side slot rows of 256), so an engine asked for 512
// runs its binary64 jobs at 256.
int block_for(int block, int precision);
// Binary64 sums in a symmetric job's source-side slots (strip rows of
// NCOMP_S x (npad - s sw) for every strip s).
size_t jpart_count(size_t n, size_t npad, size_t sw);
// Binary64 sums in a job's target-side partial rows (symmetric: triangular,
// sym_part_offset; one-sided: nchunk x NCOMP x npad).
size_t part_count(size_t n, size_t npad, size_t L, size_t tw, bool sym);

struct Diagnostics {
    double omega[3], impulse[3], angular[3];
    double ke, enstrophy, helicity;
};

// One simulation resident on the device.
struct Job {
    int64_t id = 0;
    size_t n = 0, npad = 0;
    int nchunk = 0, chunk_len = 0, ntile = 0;
    JobParams par;
    int kind = 0;  // smoothing kernel of the engine that owns the job
    bool uniform = false;
    double sigma = 0.0;
    DevBuf rec, soa, part, vol;  // soa holds x0, xs, k1, k23, rate
    int slot = -1;               // the job's entry in the engine's flag array
    int *flag = nullptr;         // device address of that entry
    DevBuf jpart;               // source-side slots of the symmetric pass
    bool sym = false;           // binary32 job on the symmetric pass
    int tw = 0;                 // targets per block of the pass
    int sw = 0;                 // targets per strip (symmetric pass)
    size_t bytes = 0;
    size_t chunk_override = 0;  // fixed chunk length (0: policy)
    JobDev dev() const;
    // factor applied to positions in the records (1/sigma for a uniform core)
    double rec_scale() const { return uniform ? 1.0 / sigma : 1.0; }
};

// Compiled kernel variant: workgroup size, targets per thread, source-loop
// unroll of the binary64 and binary32 pair kernels.
struct Variant {
    int block = 128, tpt = 1, unroll64 = 2, unroll32 = 4;
    int sym32 = 1;  // binary32 jobs: 1 symmetric pass, 0 one-sided pass
    int sym64 = 0;  // binary64 jobs: 1 symmetric pass, 0 one-sided pass (the plain method)
    int tpt32 = 0;  // targets per thread of binary32 jobs; 0: tpt
};
bool variant_exists(int kind, int tb, int tpt, int unr64, int unr32, int sym32 = 1,
                    int sym64 = 0, int tpt32 = 0);

// One wake's binary64 rates without allocating device memory (every
// intermediate in a fixed per-device scratch array of this library; all
// pointers are device arrays; asynchronous on `stream`).  At most
// noalloc_max_n() particles.  Calls that overlap on one device race.
void rates_noalloc(int kind, double nu, const double uinf[3], size_t n, const double *pos,
                   const double *vort, const double *radius, const double *volume, double *vel,
                   double *dvort, hipStream_t stream);
size_t noalloc_max_n();
size_t noalloc_scratch_bytes();

// One job's share of a server pass: RK4 stages first_stage .. first_stage +
// stages - 1 (0-based, within one step).
// One job's share of a pass: stages first_stage .. first_stage + stages - 1
// of its current RK4 step, or (parts > 1, one stage) slice `part` of that
// stage's pair work items: the slices of a stage run in consecutive passes,
// in order, and the last one runs the stage's reduction and RK4 update.
// Which items a slice holds depends only on the job and `parts`, and every
// item writes its own slots, so the result is the same however a stage is
// sliced.
struct PassJob {
    Job *job;
    int first_stage, stages;
    int part = 0, parts = 1;
};

class Engine {
  public:
    Engine(int device, int kind, const Variant &v);
    ~Engine();
    Engine(const Engine &) = delete;
    Engine &operator=(const Engine &) = delete;

    int device() const { return device_; }
    int kind() const { return kind_; }
    int block() const { return tb_; }
    Variant variant() const { return Variant{tb32_, tpt_, unr64_, unr32_, sym32_, sym64_, tpt32_}; }
    hipStream_t stream() const { return stream_; }

    // Create a job from N x 3 positions/vorticities and N radius/volume.
    // chunk_override > 0 fixes the source chunk length (multiple of block).
    // async_upload: copy and initialise on the engine's upload stream, so
    // that kernels already queued on the compute stream keep running (the
    // call returns once the job is ready).
    Job *add_job(int64_t id, size_t n, const double *pos, const double *vort,
                 const double *radius, const double *volume, const JobParams &par,
                 size_t chunk_override = 0, bool async_upload = false);
    void remove_job(int64_t id);
    // Append m particles to a job (particle emission), between steps.  The
    // device arrays grow geometrically; existing state is kept on the device.
    void append(Job *j, size_t m, const double *pos, const double *vort, const double *radius,
                const double *volume);
    Job *find(int64_t id);
    const std::vector<std::unique_ptr<Job>> &jobs() const { return jobs_; }
    void set_params(Job *j, const JobParams &par);

    // Evaluate rates of every job (solve()).
    void solve();
    // nsteps RK4 / Euler steps of every job.  use_graph replays a captured
    // step (RK4 only); timing of individual pair launches is only kept when
    // use_graph is false.
    void rk4(size_t nsteps, bool use_graph);
    void euler(size_t nsteps);

    // Host <-> device transfers (synchronous on the engine stream).
    void upload_state_soa(Job *j, const double *pos, const double *vort);  // N x 3 each
    // the same on stream s through the device staging buffer `stage`
    void upload_soa(Job *j, const double *pos, const double *vort, hipStream_t s, DevBuf &stage);
    void download_aos(Job *j, int which, double *out);  // which: 0 pos 1 vort 2 vel 3 dvort
    // particles [first, first + count) only
    void download_aos(Job *j, int which, double *out, size_t first, size_t count);
    // Queue a copy of the job's positions then vorticities (N x 3 each,
    // row-major) into device memory dst (6N doubles) on the engine stream,
    // after every step already queued; returns without waiting.
    void snapshot_aos(Job *j, double *dst);
    void upload_state_layout(Job *j, const double *state, const std::vector<int64_t> &woff);
    void download_state_layout(Job *j, int which, double *state,
                               const std::vector<int64_t> &woff);  // which: 0 state 1 rates
    // diagnostics of one job (interaction.cpp:240-383); the quadratic ones
    // (kinetic energy, enstrophy, helicity) are O(N^2) and optional
    void diagnostics(Job *j, Diagnostics *out, bool quadratic);
    // vorticity field at every particle (influence(), interaction.cpp:110-155), N x 3
    void influence(Job *j, double *field_aos);
    // non-finite flags, read back and cleared: one job's, or every job's in
    // jobs() order with one copy
    bool take_flag(Job *j);
    std::vector<char> take_flags();

    // A server pass: every listed job runs its stages, the jobs running
    // round r (0-based) being those with stages > r, each round one pair
    // launch per precision and one combine.  pass_begin queues the pass (and
    // a copy of the flags) and returns; pass_end waits for it and gives each
    // listed job's non-finite flag, cleared.  Timing as for rk4().
    void pass_begin(const std::vector<PassJob> &pj, bool use_graph);
    void pass_end(std::vector<char> *nonfinite);

    void sync();

    // timing of the last solve/rk4/euler call
    double last_total_ms = 0.0, last_pair_ms = 0.0, last_pairs = 0.0;
    int64_t pair_launches = 0;

  private:
    void build_plan();
    void capture_pass(int rounds);
    void launch_round(int r, bool timed);
    hipStream_t upload_stream();
    void launch_stage(int mode, bool timed);
    void capture_graph();
    double pairs_per_rhs() const;

    int device_, kind_, tb_, tpt_, unr64_, unr32_, sym32_, sym64_;
    int tb32_;  // threads per workgroup of binary32 jobs (tb_: binary64 jobs)
    int tb_of(int precision) const { return precision == PREC_FP32C ? tb32_ : tb_; }
    int tpt32_;  // targets per thread of binary32 jobs (tpt_: binary64 jobs)
    int tpt_of(int precision) const { return precision == PREC_FP32C ? tpt32_ : tpt_; }
    void alloc_part(Job *j);
    void (*fn64_)(const JobDev *, const Item *, const int *, int *) = nullptr;
    void (*fn32_)(const JobDev *, const Item *, const int *, int *) = nullptr;
    hipStream_t stream_ = nullptr;
    int cus_ = 0;
    int grid64_ = 0, grid32_ = 0, gridc_ = 0;
    // the grids a step of the current plan launches: no more blocks than
    // items (each block of a persistent kernel takes one queue atomic at
    // least: a few thousand idle blocks on one counter cost tens of us)
    int lgrid64_ = 1, lgrid32_ = 1, lgridc_ = 1;
    std::vector<std::unique_ptr<Job>> jobs_;
    bool plan_dirty_ = true;
    // plan tables: JobDev[njobs] | Item64[...] | Item32[...] | CItem[...] | ints
    DevBuf plan_;
    HostBuf plan_host_;
    JobDev *d_jobs_ = nullptr;
    Item *d_items64_ = nullptr, *d_items32_ = nullptr;
    CItem *d_citems_ = nullptr;
    int *d_counts_ = nullptr;  // [0] n64 [1] n32 [2] nci
    DevBuf counters_;          // [0] queue64 [1] queue32
    int n64_ = 0, n32_ = 0, nci_ = 0;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;
    // plan-table layout: region capacities and offsets, and the layout a
    // captured graph was recorded against
    size_t cap_jobs_ = 0, cap_i64_ = 0, cap_i32_ = 0, cap_ci_ = 0;
    size_t o_i64_ = 0, o_i32_ = 0, o_ci_ = 0, o_cnt_ = 0;
    uint64_t layout_version_ = 0, graph_layout_ = ~0ull;
    std::vector<hipEvent_t> ev_pool_;
    hipEvent_t ev_begin_ = nullptr, ev_end_ = nullptr;
    HostBuf stage_;
    DevBuf dstage_;
    // non-finite flags: one int per job slot
    DevBuf flags_;
    HostBuf flags_host_;
    std::vector<int> free_slots_;
    // uploads of jobs admitted while a pass runs
    hipStream_t upload_ = nullptr;
    DevBuf ustage_;
    // server pass plan: per round a JobDev table (with the round's stage per
    // job), pair items per precision, combine items and counts; regions only
    // grow (graphs hold their addresses)
    DevBuf pplan_;
    HostBuf pplan_host_;
    size_t pcap_jobs_ = 0, pcap_i64_ = 0, pcap_i32_ = 0, pcap_ci_ = 0;
    size_t p_round_ = 0, po_i64_ = 0, po_i32_ = 0, po_ci_ = 0, po_cnt_ = 0;
    uint64_t playout_ = 0;
    hipGraph_t pgraph_[4] = {nullptr, nullptr, nullptr, nullptr};
    hipGraphExec_t pgexec_[4] = {nullptr, nullptr, nullptr, nullptr};
    uint64_t pgraph_layout_[4] = {~0ull, ~0ull, ~0ull, ~0ull};
    int pn64_[4] = {0, 0, 0, 0}, pn32_[4] = {0, 0, 0, 0};
    std::vector<PassJob> cur_pass_;
    int flag_hi_ = 0;  // flag entries copied by the last pass
    double pass_pairs_ = 0.0;
    bool pass_timed_ = false;
};

}  // namespace pw
