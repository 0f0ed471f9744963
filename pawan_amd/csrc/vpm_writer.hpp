// vpm_writer.hpp -- double-buffered asynchronous .wake frame writer.
//
// __integration::integrate (src/integration/integration.cpp:16-35) writes the
// time and a frame (__interaction::write, interaction.cpp:197-202 ->
// __wake::write, wake.cpp:68-73) after every step, so the stepping waits for
// a device-to-host copy and a file write each step.  Here a frame is taken in
// three stages that overlap the next steps:
//   1. on the engine stream, after every step already queued, a kernel copies
//      positions and vorticities into the slot's device buffer (so the next
//      step may overwrite the state at once);
//   2. a copy stream waits for that and moves the buffer into the slot's
//      pinned host buffer;
//   3. a writer thread waits for the copy and appends the frame to the file.
// Two slots: a third frame waits until the older slot's file write is done.
// Frames reach the file in submission order (one thread, FIFO).
#pragma once
#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstdio>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "vpm_engine.hpp"

namespace pw {

class FrameWriter {
  public:
    explicit FrameWriter(int device);
    ~FrameWriter();  // waits for every queued frame
    FrameWriter(const FrameWriter &) = delete;
    FrameWriter &operator=(const FrameWriter &) = delete;

    // Queue a frame of job j's current state: the time t first when
    // with_time, then nWake, and per wake N, N x 3 positions, N x 3
    // vorticities, N radii (wake_n gives the wakes' sizes in order, rad the
    // radii of all particles).  Returns once the frame is queued.
    void submit(Engine *eng, Job *j, FILE *f, double t, bool with_time,
                const std::vector<size_t> &wake_n, const std::vector<double> &rad);
    // Wait until every queued frame is in its file; then report (and clear)
    // the first failure since the last call: PW_OK or an error code, with
    // its message in *msg.
    int drain(std::string *msg);
    // The first failure so far without waiting or clearing it.
    int status(std::string *msg);
    // Wait until every queued frame is in its file; failures are kept.
    void quiesce();

  private:
    struct Slot {
        DevBuf dev;               // 6N doubles: positions then vorticities
        double *host = nullptr;   // pinned copy of dev
        size_t cap = 0;           // doubles in dev and host
        hipEvent_t taken = nullptr, copied = nullptr;
        bool busy = false;        // queued or being written
        FILE *f = nullptr;
        double t = 0.0;
        bool with_time = false;
        size_t n = 0;
        std::vector<size_t> wake_n;
        std::vector<double> rad;
    };
    void run();
    void write_frame(Slot &s);

    int device_;
    Slot slot_[2];
    int next_ = 0;
    hipStream_t copy_ = nullptr;
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<int> queue_;
    bool stop_ = false;
    int err_ = 0;
    std::string msg_;
    std::thread thread_;
};

}  // namespace pw
