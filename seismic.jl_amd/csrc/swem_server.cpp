// swem_server.cpp -- serving loop: continuous batching of shot-migration requests on one GPU.
//
// Requests of any shape (2-D / 3-D, aperture, nt, band) are submitted from any thread; their
// inputs are copied, so the caller's buffers are free on return.  One worker thread owns the
// device side:
//   admission   a request is admitted when its whole device footprint (tables, work buffers,
//               records, image) fits in the free part of the server's device arena and a
//               pinned staging block fits in the pinned arena; requests behind a head that does
//               not fit may go first while the head has waited less than the skip window.
//   execution   an admitted request gets its own plan (velocity tables built on the GPU, the
//               Fourier bases of its band shared through a cache that keeps them resident),
//               its uploads, migration and image download enqueued on one of the server's
//               streams; requests on different streams overlap on the device.
//   completion  the worker polls each in-flight request's event, copies the image out of pinned
//               memory and returns the memory to the arenas, admitting the next requests.
// A request's result is computed by its own plan with deterministic kernels (no atomics, fixed
// summation order), so it is bit-identical alone and in any mix.
// Memory: one hipMalloc (device arena) and one hipHostMalloc (pinned arena) at creation; the
// loop itself allocates no device memory except the band bases (once per distinct band).
#include <algorithm>
#include <chrono>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "swem_plan.hpp"

namespace swem {
int build_bases(const swem_geometry& g, const Sampling& sm, int nf, Bases& b);
}

using namespace swem;

namespace {

constexpr size_t kAlign = 256;

size_t align_up(size_t b) { return (b + kAlign - 1) & ~(kAlign - 1); }

// First-fit range allocator with coalescing (host bookkeeping of an arena).
class RangeAlloc {
public:
    void init(size_t cap)
    {
        cap_ = cap;
        free_.clear();
        if (cap) free_[0] = cap;
        used_ = peak_ = 0;
    }
    bool alloc(size_t bytes, size_t& off)
    {
        bytes = align_up(std::max<size_t>(bytes, 1));
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                off = it->first;
                const size_t rest = it->second - bytes;
                free_.erase(it);
                if (rest) free_[off + bytes] = rest;
                used_ += bytes;
                peak_ = std::max(peak_, used_);
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t bytes)
    {
        bytes = align_up(std::max<size_t>(bytes, 1));
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t capacity() const { return cap_; }
    size_t used() const { return used_; }
    size_t peak() const { return peak_; }

private:
    size_t cap_ = 0, used_ = 0, peak_ = 0;
    std::map<size_t, size_t> free_;
};

// Bump allocator over one request's block of the device arena.
struct BumpAlloc : DevAllocator {
    char* base = nullptr;
    size_t cap = 0, used = 0;
    void* alloc(size_t bytes) override
    {
        const size_t b = align_up(bytes);
        if (used + b > cap) return nullptr;
        void* p = base + used;
        used += b;
        return p;
    }
    void release(void*, size_t) override {}
};

// Uploads of one request: copied into its pinned block, then an async copy on its stream.
struct PinnedStager : HostStager {
    char* base = nullptr;
    size_t cap = 0, used = 0;
    hipStream_t st = nullptr;
    void* take(size_t bytes)
    {
        const size_t b = align_up(bytes);
        if (used + b > cap) return nullptr;
        void* p = base + used;
        used += b;
        return p;
    }
    int put(void* dev, const void* host, size_t bytes) override
    {
        void* h = take(bytes);
        if (!h) return set_error(SWEM_ENOMEM, "request staging block exhausted (%zu bytes more)", bytes);
        std::memcpy(h, host, bytes);
        SWEM_HIP(hipMemcpyAsync(dev, h, bytes, hipMemcpyHostToDevice, st));
        return SWEM_OK;
    }
};

struct Request {
    long long id = 0;
    swem_geometry g{};
    swem_options opt{};
    std::vector<float> vel, wav, d, sx, sy;
    int nshot = 0;
    int state = SWEM_REQ_PENDING;
    int error = SWEM_OK;
    bool claimed = false;  // a thread waits in swem_batch_wait (one waiter per request)
    // fusion: single-shot requests with equal geometry, options, velocity and wavelet (equal
    // key, then compared in full) run as the shots of one launch on one plan
    uint64_t key = 0;                 // 0: not fusable (several shots)
    std::vector<Request*> group;      // admitted: the requests of this launch, itself first
    std::string message;
    double t_submit = 0, t_admit = 0, t_done = 0;
    // resources while admitted
    size_t dev_off = 0, dev_bytes = 0, pin_off = 0, pin_bytes = 0;
    swem_plan* plan = nullptr;
    hipEvent_t done = nullptr;
    int stream = -1;
    float* pin_out = nullptr;
    std::atomic<bool> enqueued{false};  // a preparation thread has enqueued the launch and its event
    // allocator of the request's device block: the plan's buffers refer to it until released
    BumpAlloc bump;
    std::vector<float> m;
};

// Fourier bases per band, resident for the server's lifetime (a handful of distinct bands).
struct ServerBases : BasisCache {
    std::map<std::vector<long long>, std::unique_ptr<Bases>> entries;
    std::mutex mu;                 // the preparation threads look up and build concurrently
    std::atomic<size_t> bytes{0};  // read by swem_server_get_stats
    const Bases* find_or_build(const swem_geometry& g, const Sampling& s, int nf) override
    {
        std::lock_guard<std::mutex> lk(mu);
        long long dtbits = 0;
        std::memcpy(&dtbits, &g.dt, sizeof(float));
        std::vector<long long> key{g.nt, dtbits, s.ntfft, s.ifmin, nf, s.fop_n};
        auto it = entries.find(key);
        if (it != entries.end()) return it->second.get();
        // built with plain device allocations and synchronous uploads, outside the request
        DevAllocator* da = tl_dev_alloc;
        HostStager* hs = tl_stager;
        tl_dev_alloc = nullptr;
        tl_stager = nullptr;
        auto b = std::make_unique<Bases>();
        const int rc = build_bases(g, s, nf, *b);
        tl_dev_alloc = da;
        tl_stager = hs;
        if (rc) return nullptr;
        bytes += b->adj.bytes() + b->fwd.bytes() + b->src.bytes();
        const Bases* out = b.get();
        entries.emplace(key, std::move(b));
        return out;
    }
    ~ServerBases() override
    {
        for (auto& e : entries) {
            e.second->adj.release();
            e.second->fwd.release();
            e.second->src.release();
        }
    }
};

double now_ms(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct swem_server {
    swem_server_config cfg{};
    char* dev_base = nullptr;
    char* pin_base = nullptr;
    RangeAlloc dev, pin;
    std::vector<hipStream_t> streams;
    std::vector<int> stream_load;
    std::vector<hipEvent_t> event_pool;
    ServerBases bases;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> pending;
    std::vector<Request*> inflight;  // worker only
    std::unordered_map<long long, std::unique_ptr<Request>> all;
    std::thread worker;
    bool stop = false;
    // preparation threads: tables, uploads and the launch of admitted requests (host work that
    // would otherwise serialise admissions on the worker)
    std::vector<std::thread> prep;
    std::deque<Request*> jobs;
    std::condition_variable cv_prep;
    bool prep_stop = false;
    long long next_id = 1;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    // statistics
    long long submitted = 0, completed = 0, failed = 0, fused_launches = 0, fused_requests = 0;
    double busy_since = -1, busy_ms = 0, first_submit = -1, last_done = 0;
};

namespace {

// Device and pinned footprint of a launch of nshot shots of r's shape, with nimg images (plan
// set up, nothing allocated yet).
void footprint(const Request& r, const swem_plan* p, int nshot, int nimg, size_t& dev_bytes, size_t& pin_bytes)
{
    const size_t nm = p->nm;
    const size_t dbytes = (size_t)nshot * nm * r.g.nt * sizeof(float), mbytes = nimg * nm * r.g.nz * sizeof(float);
    dev_bytes = plan_migrate_bytes(p, nshot) + align_up(dbytes) + align_up(mbytes) + (1 << 20);
    pin_bytes = plan_upload_bytes(p, nshot) + align_up(dbytes) + align_up(mbytes) + (1 << 16);
}

// Fusion key of a single-shot request: a 64-bit hash of its geometry, options and wavelet and of
// 64 velocity samples spread over the model, so that a submission does not read a 3-D velocity
// through (equal keys are confirmed by a full comparison before two requests are fused).
uint64_t fusion_key(const Request& r)
{
    uint64_t h = 0x9e3779b97f4a7c15ull;
    auto mix = [&h](const void* data, size_t bytes) {
        const unsigned char* b = static_cast<const unsigned char*>(data);
        size_t i = 0;
        for (; i + 8 <= bytes; i += 8) {
            uint64_t w;
            std::memcpy(&w, b + i, 8);
            h = (h ^ w) * 0x100000001b3ull;
            h ^= h >> 29;
        }
        for (; i < bytes; i++) h = (h ^ b[i]) * 0x100000001b3ull;
    };
    mix(&r.g, sizeof r.g);
    mix(&r.opt.pspi_mode, sizeof r.opt.pspi_mode);
    mix(&r.opt.precision, sizeof r.opt.precision);
    const size_t nv = r.vel.size(), step = std::max<size_t>(1, nv / 64);
    for (size_t i = 0; i < nv; i += step) mix(&r.vel[i], sizeof(float));
    mix(&r.vel[nv - 1], sizeof(float));
    mix(r.wav.data(), r.wav.size() * sizeof(float));
    return h ? h : 1;
}

bool fusable(const Request& a, const Request& b)
{
    return a.key && a.key == b.key && a.nshot == 1 && b.nshot == 1 &&
           std::memcmp(&a.g, &b.g, sizeof a.g) == 0 && a.opt.pspi_mode == b.opt.pspi_mode &&
           a.opt.precision == b.opt.precision && a.vel == b.vel && a.wav == b.wav;
}

// Frees what an admitted request holds (its work has completed or never started).  Worker only.
void release(swem_server* s, Request* r)
{
    if (r->plan) {
        plan_free(r->plan);
        r->plan = nullptr;
    }
    if (r->dev_bytes) s->dev.release(r->dev_off, r->dev_bytes);
    if (r->pin_bytes) s->pin.release(r->pin_off, r->pin_bytes);
    r->dev_bytes = r->pin_bytes = 0;
    if (r->done) s->event_pool.push_back(r->done);
    r->done = nullptr;
    if (r->stream >= 0) s->stream_load[r->stream]--;
    r->stream = -1;
}

enum AdmitResult { kAdmitted, kNoFit, kFailed };

// Tries to admit the launch of grp (worker thread, server lock not held): grp[0] alone, or
// several fusable single-shot requests as the shots of one launch with one image each.  kAdmitted
// reserves the plan (set up, nothing built), the arena blocks, a stream and an event; a preparation
// thread then builds and enqueues the launch (enqueue_launch).  kFailed leaves the status in
// code / msg.
AdmitResult admit(swem_server* s, const std::vector<Request*>& grp, int& code, std::string& msg)
{
    Request* r = grp[0];
    const bool fused = grp.size() > 1;
    const int nshot = fused ? (int)grp.size() : r->nshot, nimg = fused ? (int)grp.size() : 1;
    code = SWEM_OK;
    swem_plan* p = new swem_plan();
    p->served = true;
    p->use_graphs = 0;  // one plan per request: a captured graph would not be replayed
    p->basis_cache = &s->bases;
    swem_options o = r->opt;
    o.device = s->cfg.device;
    o.scratch_bytes = s->dev.capacity() / 2;
    o.table_bytes = s->dev.capacity() / 4;
    int rc = plan_setup(p, &r->g, r->vel.data(), &o);
    if (rc) {
        code = rc;
        msg = swem_last_error();
        plan_free(p);
        return kFailed;
    }
    size_t dev_bytes, pin_bytes;
    footprint(*r, p, nshot, nimg, dev_bytes, pin_bytes);
    if (dev_bytes > s->dev.capacity() || pin_bytes > s->pin.capacity()) {
        plan_free(p);
        code = SWEM_ENOMEM;
        msg = "request larger than the server's arenas";
        return kFailed;
    }
    size_t doff, poff;
    if (!s->dev.alloc(dev_bytes, doff)) {
        plan_free(p);
        return kNoFit;
    }
    if (!s->pin.alloc(pin_bytes, poff)) {
        s->dev.release(doff, dev_bytes);
        plan_free(p);
        return kNoFit;
    }
    r->plan = p;
    r->group = grp;
    r->dev_off = doff;
    r->dev_bytes = dev_bytes;
    r->pin_off = poff;
    r->pin_bytes = pin_bytes;
    int si = 0;
    for (int i = 1; i < (int)s->streams.size(); i++)
        if (s->stream_load[i] < s->stream_load[si]) si = i;
    r->stream = si;
    s->stream_load[si]++;
    hipStream_t st = s->streams[si];
    p->build_stream = st;
    if (s->event_pool.empty()) {
        hipEvent_t e;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            release(s, r);
            code = SWEM_EDEVICE;
            msg = "hipEventCreate failed";
            return kFailed;
        }
        s->event_pool.push_back(e);
    }
    r->done = s->event_pool.back();
    s->event_pool.pop_back();

    return kAdmitted;
}

// Builds the plan of an admitted launch (its tables into the request's arena block), stages the
// records, enqueues the migration and the image download on the request's stream and records its
// event (a preparation thread).  Returns the status of the enqueue; the event is recorded either
// way, so the arena blocks are released only after the stream has passed whatever was enqueued.
int enqueue_launch(swem_server* s, Request* r, std::string& msg)
{
    const std::vector<Request*>& grp = r->group;
    const bool fused = grp.size() > 1;
    const int nshot = fused ? (int)grp.size() : r->nshot, nimg = fused ? (int)grp.size() : 1;
    swem_plan* p = r->plan;
    hipStream_t st = s->streams[r->stream];
    BumpAlloc& bump = r->bump;
    bump.base = s->dev_base + r->dev_off;
    bump.cap = r->dev_bytes;
    bump.used = 0;
    PinnedStager stager;
    stager.base = s->pin_base + r->pin_off;
    stager.cap = r->pin_bytes;
    stager.st = st;
    tl_dev_alloc = &bump;
    tl_stager = &stager;
    int rc = SWEM_OK;
    const size_t nm = p->nm, rec = nm * r->g.nt;
    DevBuf<float> d_dev, m_dev;
    std::vector<float> sx, sy;
    for (Request* q : grp) {
        sx.insert(sx.end(), q->sx.begin(), q->sx.end());
        sy.insert(sy.end(), q->sy.begin(), q->sy.end());
    }
    rc = plan_build(p, r->wav.data());
    if (!rc) rc = d_dev.alloc((size_t)nshot * rec);
    for (size_t i = 0; i < grp.size() && !rc; i++)
        rc = stager.put(d_dev.p + (fused ? i * rec : 0), grp[i]->d.data(), grp[i]->d.size() * sizeof(float));
    if (!rc) rc = m_dev.alloc((size_t)nimg * nm * r->g.nz);
    if (!rc) rc = plan_migrate(p, d_dev.p, sx.data(), sy.data(), nshot, m_dev.p, 0, fused ? 1 : 0, st);
    if (!rc) {
        r->pin_out = static_cast<float*>(stager.take(m_dev.bytes()));
        if (!r->pin_out) rc = set_error(SWEM_ENOMEM, "request staging block exhausted (image)");
    }
    if (!rc && hipMemcpyAsync(r->pin_out, m_dev.p, m_dev.bytes(), hipMemcpyDeviceToHost, st) != hipSuccess)
        rc = set_error(SWEM_EDEVICE, "image download could not be enqueued");
    tl_dev_alloc = nullptr;
    tl_stager = nullptr;
    d_dev.p = m_dev.p = nullptr;  // both live in the request block
    if (rc) msg = swem_last_error();
    (void)hipEventRecord(r->done, st);
    return rc;
}

void prep_main(swem_server* s)
{
    (void)hipSetDevice(s->cfg.device);
    for (;;) {
        Request* r = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_prep.wait(lk, [s] { return s->prep_stop || !s->jobs.empty(); });
            if (s->jobs.empty()) return;
            r = s->jobs.front();
            s->jobs.pop_front();
        }
        std::string msg;
        const int rc = enqueue_launch(s, r, msg);
        if (rc) {
            std::lock_guard<std::mutex> lk(s->mu);
            r->error = rc;
            r->message = msg;
        }
        r->enqueued.store(true, std::memory_order_release);
    }
}

void worker_main(swem_server* s)
{
    (void)hipSetDevice(s->cfg.device);
    std::vector<Request*> cand;
    for (;;) {
        bool progress = false;
        // completions
        for (size_t i = 0; i < s->inflight.size();) {
            Request* r = s->inflight[i];
            if (!r->enqueued.load(std::memory_order_acquire)) {  // still being prepared
                ++i;
                continue;
            }
            const hipError_t q = hipEventQuery(r->done);
            if (q == hipErrorNotReady) {
                ++i;
                continue;
            }
            {
                std::lock_guard<std::mutex> lk(s->mu);
                if (q != hipSuccess && !r->error) {
                    r->error = SWEM_EDEVICE;
                    r->message = hipGetErrorString(q);
                }
                const double t = now_ms(s->t0);
                const size_t nimg = (size_t)r->g.nmx * r->g.nmy * r->g.nz;
                // every request of the launch (the leader last: release() uses it)
                for (size_t k = r->group.size(); k-- > 0;) {
                    Request* m = r->group[k];
                    if (m != r) {
                        m->error = r->error;
                        m->message = r->message;
                    }
                    if (!r->error) {
                        m->m.assign(r->pin_out + k * nimg, r->pin_out + (k + 1) * nimg);
                        m->state = SWEM_REQ_DONE;
                        s->completed++;
                    } else {
                        m->state = SWEM_REQ_FAILED;
                        s->failed++;
                    }
                    m->t_done = t;
                }
                release(s, r);
                r->group.clear();
                s->last_done = t;
                s->inflight.erase(s->inflight.begin() + i);
                if (s->inflight.empty() && s->busy_since >= 0) {
                    s->busy_ms += t - s->busy_since;
                    s->busy_since = -1;
                }
            }
            s->cv_done.notify_all();
            progress = true;
        }
        // admissions: candidates from the head of the queue, admitted without the lock held
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (s->stop && s->pending.empty() && s->inflight.empty()) break;
            if (!s->pending.empty() && (int)s->inflight.size() < s->cfg.max_inflight) {
                cand.assign(s->pending.begin(), s->pending.begin() + std::min<size_t>(16, s->pending.size()));
            } else {
                cand.clear();
                if (s->inflight.empty()) {
                    s->cv_work.wait(lk, [s] { return s->stop || !s->pending.empty(); });
                    continue;
                }
            }
        }
        std::vector<Request*> grp;
        for (size_t k = 0; k < cand.size(); k++) {
            Request* r = cand[k];
            if ((int)s->inflight.size() >= s->cfg.max_inflight) break;
            // the pending requests this one fuses with (continuous batching: those already queued)
            grp.assign(1, r);
            {
                std::lock_guard<std::mutex> lk(s->mu);
                if (r->state != SWEM_REQ_PENDING) continue;  // taken into an earlier launch
                if (r->key && s->cfg.max_fuse > 1)
                    for (Request* q : s->pending) {
                        if ((int)grp.size() >= s->cfg.max_fuse) break;
                        if (q != r && q->state == SWEM_REQ_PENDING && q->key == r->key) grp.push_back(q);
                    }
            }
            // equal keys confirmed in full outside the lock: a request's inputs do not change after
            // submission, and only this thread takes requests out of the pending state
            grp.erase(std::remove_if(grp.begin() + 1, grp.end(), [r](const Request* q) { return !fusable(*r, *q); }),
                      grp.end());
            int code = 0;
            std::string msg;
            AdmitResult res = admit(s, grp, code, msg);
            if (res == kNoFit && grp.size() > 1) {  // the launch of the group does not fit: r alone
                grp.resize(1);
                res = admit(s, grp, code, msg);
            }
            std::unique_lock<std::mutex> lk(s->mu);
            const double t = now_ms(s->t0);
            if (res == kNoFit) {
                if (!s->inflight.empty()) {
                    // head of line does not fit: smaller requests may pass while it is young
                    if (k == 0 && t - r->t_submit > s->cfg.skip_window_ms) break;
                    continue;
                }
                code = SWEM_ENOMEM;  // does not fit even on an idle device
                msg = "request does not fit the server's free arenas";
            }
            for (Request* m : grp) {
                s->pending.erase(std::find(s->pending.begin(), s->pending.end(), m));
                m->error = code;
                m->message = msg;
                if (res == kAdmitted) {
                    m->state = SWEM_REQ_RUNNING;
                    m->t_admit = t;
                } else {
                    m->state = SWEM_REQ_FAILED;
                    m->t_done = t;
                    s->failed++;
                }
            }
            if (res == kAdmitted) {
                if (s->inflight.empty()) s->busy_since = t;
                s->inflight.push_back(r);
                if (grp.size() > 1) {
                    s->fused_launches++;
                    s->fused_requests += (long long)grp.size();
                }
                s->jobs.push_back(r);
                s->cv_prep.notify_one();
            } else {
                lk.unlock();
                s->cv_done.notify_all();
            }
            progress = true;
        }
        if (!progress) std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
}

}  // namespace

extern "C" {

void swem_server_config_init(swem_server_config* c)
{
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->streams = 4;
    c->max_inflight = 8;
    c->max_pending = 4096;
    c->skip_window_ms = 500.0;
    c->max_fuse = 32;
}

int swem_server_create(const swem_server_config* cfg, swem_server** out)
{
    if (!out) return set_error(SWEM_EINVAL, "server_create: null argument");
    *out = nullptr;
    swem_server_config c;
    swem_server_config_init(&c);
    if (cfg) c = *cfg;
    if (c.streams < 1 || c.streams > 16 || c.max_inflight < 1 || c.max_pending < 1 || c.max_fuse < 1 ||
        c.max_fuse > 1024)
        return set_error(SWEM_EINVAL,
                         "server config: streams 1..16, max_inflight >= 1, max_pending >= 1, max_fuse 1..1024");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || c.device < 0 || c.device >= count)
        return set_error(SWEM_EDEVICE, "server: device %d not available", c.device);
    SWEM_HIP(hipSetDevice(c.device));
    if (!c.device_bytes) {
        size_t fr = 0, tot = 0;
        SWEM_HIP(hipMemGetInfo(&fr, &tot));
        c.device_bytes = std::min((size_t)64 << 30, fr / 2);
    }
    if (!c.pinned_bytes) c.pinned_bytes = (size_t)4 << 30;
    auto s = std::make_unique<swem_server>();
    s->cfg = c;
    SWEM_HIP(hipMalloc((void**)&s->dev_base, c.device_bytes));
    if (hipHostMalloc((void**)&s->pin_base, c.pinned_bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipFree(s->dev_base);
        return set_error(SWEM_ENOMEM, "server: %zu bytes of pinned host memory", c.pinned_bytes);
    }
    s->dev.init(c.device_bytes);
    s->pin.init(c.pinned_bytes);
    for (int i = 0; i < c.streams; i++) {
        hipStream_t st;
        SWEM_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        s->streams.push_back(st);
        s->stream_load.push_back(0);
    }
    swem_server* raw = s.release();
    for (int i = 0; i < c.streams; i++) raw->prep.emplace_back(prep_main, raw);
    raw->worker = std::thread(worker_main, raw);
    *out = raw;
    return SWEM_OK;
}

int swem_server_destroy(swem_server* s)
{
    if (!s) return SWEM_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();  // returns when nothing is pending or in flight
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->prep_stop = true;
    }
    s->cv_prep.notify_all();
    for (std::thread& t : s->prep) t.join();
    (void)hipSetDevice(s->cfg.device);
    for (hipStream_t st : s->streams) (void)hipStreamSynchronize(st);
    for (hipStream_t st : s->streams) (void)hipStreamDestroy(st);
    for (hipEvent_t e : s->event_pool) (void)hipEventDestroy(e);
    (void)hipFree(s->dev_base);
    (void)hipHostFree(s->pin_base);
    delete s;
    return SWEM_OK;
}

int swem_batch_submit(swem_server* s, const swem_request* q, long long* id)
{
    if (!s || !q || !id || !q->vel || !q->wav || !q->d || !q->sx || q->nshot < 1)
        return set_error(SWEM_EINVAL, "batch_submit: null argument or nshot < 1");
    const swem_geometry& g = q->g;
    if (g.nt < 1 || g.nmx < 1 || g.nmy < 1 || g.nz < 1) return set_error(SWEM_EINVAL, "batch_submit: empty geometry");
    auto r = std::make_unique<Request>();
    r->g = g;
    swem_options_init(&r->opt);
    r->opt.pspi_mode = q->pspi_mode;
    r->opt.precision = q->precision;
    const size_t nm = (size_t)g.nmx * g.nmy;
    r->vel.assign(q->vel, q->vel + nm * g.nz);
    r->wav.assign(q->wav, q->wav + g.nt);
    r->d.assign(q->d, q->d + (size_t)q->nshot * nm * g.nt);
    r->sx.assign(q->sx, q->sx + q->nshot);
    r->sy.assign(q->nshot, 0.f);
    if (q->sy) r->sy.assign(q->sy, q->sy + q->nshot);
    r->nshot = q->nshot;
    if (r->nshot == 1) r->key = fusion_key(*r);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if ((int)s->pending.size() >= s->cfg.max_pending) return set_error(SWEM_EBUSY, "serving queue full");
        r->id = s->next_id++;
        r->t_submit = now_ms(s->t0);
        if (s->first_submit < 0) s->first_submit = r->t_submit;
        *id = r->id;
        s->pending.push_back(r.get());
        s->all.emplace(r->id, std::move(r));
        s->submitted++;
    }
    s->cv_work.notify_one();
    return SWEM_OK;
}

static void fill_status(const Request* r, swem_request_status* st)
{
    if (!st) return;
    std::memset(st, 0, sizeof(*st));
    st->state = r->state;
    st->error = r->error;
    st->submit_ms = r->t_submit;
    st->admit_ms = r->t_admit;
    st->done_ms = r->t_done;
    std::strncpy(st->message, r->message.c_str(), sizeof(st->message) - 1);
}

int swem_batch_poll(swem_server* s, long long id, swem_request_status* st)
{
    if (!s || !st) return set_error(SWEM_EINVAL, "batch_poll: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) return set_error(SWEM_EINVAL, "batch_poll: unknown request %lld", id);
    fill_status(it->second.get(), st);
    return SWEM_OK;
}

int swem_batch_wait(swem_server* s, long long id, float* m, swem_request_status* st)
{
    if (!s) return set_error(SWEM_EINVAL, "batch_wait: null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) return set_error(SWEM_EINVAL, "batch_wait: unknown request %lld", id);
    Request* r = it->second.get();
    // the waiter forgets the request when it wakes, so a second waiter would be left holding a
    // freed request: refuse it
    if (r->claimed) return set_error(SWEM_EINVAL, "batch_wait: request %lld already has a waiter", id);
    r->claimed = true;
    s->cv_done.wait(lk, [r] { return r->state == SWEM_REQ_DONE || r->state == SWEM_REQ_FAILED; });
    fill_status(r, st);
    int rc = SWEM_OK;
    if (r->state == SWEM_REQ_DONE) {
        if (m) std::memcpy(m, r->m.data(), r->m.size() * sizeof(float));
    } else {
        rc = set_error(r->error ? r->error : SWEM_EDEVICE, "request %lld failed: %s", id, r->message.c_str());
    }
    s->all.erase(it);
    return rc;
}

int swem_server_get_stats(swem_server* s, swem_server_stats* out)
{
    if (!s || !out) return set_error(SWEM_EINVAL, "server_stats: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(out, 0, sizeof(*out));
    out->submitted = s->submitted;
    out->completed = s->completed;
    out->failed = s->failed;
    out->pending = (int)s->pending.size();
    out->inflight = (int)s->inflight.size();
    out->device_bytes = s->dev.capacity();
    out->device_bytes_used = s->dev.used();
    out->device_bytes_peak = s->dev.peak();
    out->busy_ms = s->busy_ms + (s->busy_since >= 0 ? now_ms(s->t0) - s->busy_since : 0.0);
    out->basis_bytes = s->bases.bytes.load();
    out->fused_launches = s->fused_launches;
    out->fused_requests = s->fused_requests;
    return SWEM_OK;
}

}  // extern "C"
