// vpm_writer.hip -- the asynchronous .wake frame writer (vpm_writer.hpp).
#include "vpm_writer.hpp"

#include "../../include/pawan_hip.h"

namespace pw {

FrameWriter::FrameWriter(int device) : device_(device) {
    PW_HIP(hipSetDevice(device_));
    PW_HIP(hipStreamCreateWithFlags(&copy_, hipStreamNonBlocking));
    for (Slot &s : slot_) {
        PW_HIP(hipEventCreateWithFlags(&s.taken, hipEventDisableTiming));
        PW_HIP(hipEventCreateWithFlags(&s.copied, hipEventDisableTiming));
    }
    thread_ = std::thread([this] { run(); });
}

FrameWriter::~FrameWriter() {
    {
        std::unique_lock<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    if (thread_.joinable()) thread_.join();  // the thread empties the queue first
    (void)hipSetDevice(device_);
    for (Slot &s : slot_) {
        if (s.host) (void)hipHostFree(s.host);
        if (s.taken) (void)hipEventDestroy(s.taken);
        if (s.copied) (void)hipEventDestroy(s.copied);
    }
    if (copy_) (void)hipStreamDestroy(copy_);
}

void FrameWriter::submit(Engine *eng, Job *j, FILE *f, double t, bool with_time,
                         const std::vector<size_t> &wake_n, const std::vector<double> &rad) {
    const int k = next_;
    Slot &s = slot_[k];
    {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return !s.busy; });
    }
    // the slot is free: its last copy has completed and its file write is done
    PW_HIP(hipSetDevice(device_));
    const size_t n = j->n, need = 6 * n;
    if (s.cap < need) {
        const size_t cap = need + need / 4;
        s.dev.reset(cap * sizeof(double));
        if (s.host) {
            (void)hipHostFree(s.host);
            s.host = nullptr;
            s.cap = 0;
        }
        PW_HIP(hipHostMalloc((void **)&s.host, cap * sizeof(double), hipHostMallocDefault));
        s.cap = cap;
    }
    if (n) {
        eng->snapshot_aos(j, s.dev.as<double>());
        PW_HIP(hipEventRecord(s.taken, eng->stream()));
        PW_HIP(hipStreamWaitEvent(copy_, s.taken, 0));
        PW_HIP(hipMemcpyAsync(s.host, s.dev.p, need * sizeof(double), hipMemcpyDeviceToHost,
                              copy_));
    }
    PW_HIP(hipEventRecord(s.copied, copy_));
    s.f = f;
    s.t = t;
    s.with_time = with_time;
    s.n = n;
    s.wake_n = wake_n;
    s.rad.assign(rad.begin(), rad.begin() + (ptrdiff_t)n);
    {
        std::unique_lock<std::mutex> lk(mu_);
        s.busy = true;
        queue_.push_back(k);
    }
    next_ ^= 1;
    cv_.notify_all();
}

void FrameWriter::write_frame(Slot &s) {
    FILE *f = s.f;
    bool ok = true;
    if (s.with_time) ok = std::fwrite(&s.t, sizeof(double), 1, f) == 1;
    const size_t nw = s.wake_n.size();
    ok = ok && std::fwrite(&nw, sizeof(size_t), 1, f) == 1;
    size_t o = 0;
    const double *pos = s.host, *vort = s.host + 3 * s.n;
    for (size_t w = 0; w < nw && ok; ++w) {
        const size_t m = s.wake_n[w];
        ok = std::fwrite(&m, sizeof(size_t), 1, f) == 1 &&
             std::fwrite(pos + 3 * o, sizeof(double), 3 * m, f) == 3 * m &&
             std::fwrite(vort + 3 * o, sizeof(double), 3 * m, f) == 3 * m &&
             std::fwrite(s.rad.data() + o, sizeof(double), m, f) == m;
        o += m;
    }
    if (!ok) throw Error(PW_EINVAL, "write to the file failed");
}

void FrameWriter::run() {
    (void)hipSetDevice(device_);
    for (;;) {
        int k;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return stop_ || !queue_.empty(); });
            if (queue_.empty()) return;  // stop_ and nothing left
            k = queue_.front();
        }
        Slot &s = slot_[k];
        int code = PW_OK;
        std::string what;
        const hipError_t e = hipEventSynchronize(s.copied);
        if (e != hipSuccess) {
            code = PW_EHIP;
            what = std::string("frame copy: ") + hipGetErrorString(e);
        } else {
            try {
                write_frame(s);
            } catch (const Error &x) {
                code = x.code;
                what = x.what();
            }
        }
        {
            std::unique_lock<std::mutex> lk(mu_);
            if (code != PW_OK && err_ == PW_OK) {
                err_ = code;
                msg_ = what;
            }
            queue_.pop_front();
            s.busy = false;
        }
        cv_.notify_all();
    }
}

int FrameWriter::drain(std::string *msg) {
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait(lk, [&] { return queue_.empty(); });
    const int code = err_;
    if (msg) *msg = msg_;
    err_ = PW_OK;
    msg_.clear();
    return code;
}

void FrameWriter::quiesce() {
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait(lk, [&] { return queue_.empty(); });
}

int FrameWriter::status(std::string *msg) {
    std::unique_lock<std::mutex> lk(mu_);
    if (msg) *msg = msg_;
    return err_;
}

}  // namespace pw
