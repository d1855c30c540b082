// Orchestration of the MI355X full-field solve; see engine.hpp. Phase order,
// decisions and counters follow src/pipeline/full_field_solver.cpp:141-586,
// full_field_anchors.cpp:283-607, full_field_path_{a,b,c}.cpp and
// full_field_solver_stats.cpp of the reference; every per-pixel and per-node
// computation runs in a kernel.
#include "engine.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <queue>
#include <thread>

#include "seeding.hpp"

namespace sdic {

namespace {

using clk = std::chrono::steady_clock;
double ms_since(clk::time_point t0) { return std::chrono::duration<double, std::milli>(clk::now() - t0).count(); }

constexpr int kErrRoi = -2, kErrInit = -3, kErrStrainWindow = -4, kErrSeed = -5, kErrCancelled = -99;

struct HipFail {
    hipError_t e;
};
inline void ck(hipError_t e) {
    if (e != hipSuccess) throw HipFail{e};
}

bool clip_axis(int& lo, int& len, int step, int extent) {
    int64_t a = lo;
    const int64_t b = std::min<int64_t>((int64_t)lo + (int64_t)len, extent);
    if (a < 0) a += ((-a + step - 1) / step) * (int64_t)step;
    if (b <= a) return false;
    lo = (int)a;
    len = (int)(b - a);
    return true;
}

// validate_params (full_field_solver.cpp:97-111)
int validate(Params& p, int width, int height) {
    if (width <= 0 || height <= 0) return kErrInit;
    if ((int64_t)width * height > (int64_t)INT32_MAX) return kErrInit;
    if (p.step <= 0) return kErrRoi;
    if (p.subset < 3 || p.subset % 2 == 0) return kErrInit;
    if (p.rect_w <= 0 || p.rect_h <= 0) return kErrRoi;
    if (!clip_axis(p.rect_x, p.rect_w, p.step, width)) return kErrRoi;
    if (!clip_axis(p.rect_y, p.rect_h, p.step, height)) return kErrRoi;
    if (p.rect_w / p.step <= 0 || p.rect_h / p.step <= 0) return kErrRoi;
    const int64_t gmin = std::min(p.rect_w / p.step, p.rect_h / p.step);
    if ((int64_t)p.strain_window > 2 * (gmin + 1) * (int64_t)p.step) return kErrInit;
    return 0;
}

// StrainCalculator::vsg_window_node_count (strain_calculator.cpp:12-25)
int vsg_count(int step, int window) {
    if (step <= 0) return 0;
    const float radius = window / 2.0f;
    const int gr = (int)std::ceil(radius / (float)step);
    const double r2 = (double)(radius * radius) + kVsgRadiusTiny;
    int c = 0;
    for (int dy = -gr; dy <= gr; ++dy)
        for (int dx = -gr; dx <= gr; ++dx) {
            const double px = (double)(dx * step), py = (double)(dy * step);
            if (px * px + py * py <= r2) c++;
        }
    return c;
}

struct Counters {   // one ThreadStats bucket
    int points = 0, iters = 0, calls = 0, saved = 0, dead = 0;
    void record(const SolveOut& o) {
        iters += o.iters;
        if (o.simplex) {
            calls++;
            if (o.status == 0 && o.corr <= kCorrAccept) saved++;
            else dead++;
        }
    }
};

// Progress ticker: 10..90 % from a thread of the solve, about every 100 ms
// (full_field_solver.cpp:334-373).
class Ticker {
public:
    Ticker(const std::function<void(int)>& cb, int total) : cb_(cb), total_(total) {
        if (cb_) th_ = std::thread([this] { loop(); });
    }
    ~Ticker() { stop(); }
    void set(int solved) { solved_.store(solved, std::memory_order_relaxed); }
    void stop() {
        {
            std::lock_guard<std::mutex> lk(m_);
            stop_ = true;
        }
        cv_.notify_all();
        if (th_.joinable()) th_.join();
    }

private:
    void loop() {
        std::unique_lock<std::mutex> lk(m_);
        while (!stop_) {
            const int pct = std::max(0, std::min(100, (int)(((float)solved_.load() / (float)std::max(1, total_)) * 80.0f + 10.0f)));
            lk.unlock();
            try {
                cb_(pct);
            } catch (...) {
            }
            lk.lock();
            cv_.wait_for(lk, std::chrono::milliseconds(100), [this] { return stop_; });
        }
    }
    std::function<void(int)> cb_;
    int total_;
    std::atomic<int> solved_{0};
    std::thread th_;
    std::mutex m_;
    std::condition_variable cv_;
    bool stop_ = false;
};

}  // namespace

Engine::Engine() {
    ck(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
    ck(hipEventCreate(&ev0_));
    ck(hipEventCreate(&ev1_));
}

Engine::~Engine() {
    if (st_) (void)hipStreamSynchronize(st_);
    for (Chain* c : {&tail_chain_, &strain_chain_})
        if (c->exec) (void)hipGraphExecDestroy(c->exec);
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    for (hipEvent_t e : class_ev_) (void)hipEventDestroy(e);
    for (hipStream_t s : class_st_) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
    }
    if (st_) (void)hipStreamDestroy(st_);
}

void Engine::reset() {
    have_ref_ = false;
    grid_valid_ = false;
    g16_fmt_ = -1;
    w_ = h_ = 0;
}

int Engine::set_reference(const uint8_t* gray, bool on_device, int w, int h, const uint8_t* mask) {
    reset();
    if (!gray || w <= 0 || h <= 0 || (int64_t)w * h > (int64_t)INT32_MAX) return kErrInit;
    try {
        const size_t n = (size_t)w * h;
        if (!ref_u8_.reserve(n) || !I_.reserve(n) || !gx_.reserve(n) || !gy_.reserve(n) || !mask_d_.reserve(n))
            return kErrInit;
        ck(hipMemcpyAsync(ref_u8_.p, gray, n, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st_));
        if (mask) ck(hipMemcpyAsync(mask_d_.p, mask, n, hipMemcpyHostToDevice, st_));
        ck(launch_prep_ref_u8(ref_u8_.p, mask ? mask_d_.p : nullptr, w, h, I_.p, gx_.p, gy_.p, st_));
        // reference pyramid of the global-motion search (levels 1..3)
        int pw = w, ph = h;
        pyr_ref_levels_ = 0;
        for (int l = 1; l <= 3; ++l) {
            const int nw = pw / 2, nh = ph / 2;
            if (nw < 1 || nh < 1) break;
            if (!pyr_r_[l].reserve((size_t)nw * nh)) return kErrInit;
            if (l == 1) ck(launch_downsample_u8(ref_u8_.p, pw, ph, pyr_r_[1].p, st_));
            else ck(launch_downsample_f32(pyr_r_[l - 1].p, pw, ph, pyr_r_[l].p, st_));
            pw = nw;
            ph = nh;
            pyr_ref_levels_ = l;
        }
        ck(hipStreamSynchronize(st_));
    } catch (const HipFail&) {
        return kErrInit;
    }
    w_ = w;
    h_ = h;
    have_ref_ = true;
    return 0;
}

// Grid scan and Hessian pool, cached while rect / step / subset / mask hold.
namespace {
const Engine::Tune kShapes[6] = {{1, 0}, {2, 0}, {4, 0}, {1, 8}, {2, 8}, {4, 8}};
int shape_count(bool listed) { return listed ? 3 : 6; }   // tiles order the whole grid only
}  // namespace

Engine::GraphStats Engine::graph_stats() const {
    GraphStats g;
    for (const Chain* c : {&tail_chain_, &strain_chain_}) {
        g.captures += c->captures;
        g.replays += c->replays;
    }
    return g;
}

// Runs a chain of stream work: eagerly, or as a replay of its captured graph.
// The first call with a key runs eagerly and keeps the key; a second call in
// a row with the same key captures the chain, and later ones replay it.
template <typename F>
void Engine::run_chain(Chain& c, const std::vector<uint64_t>& key, F&& enqueue) {
    if (!use_graphs) {
        enqueue();
        return;
    }
    if (c.exec && c.key == key) {
        ck(hipGraphLaunch(c.exec, st_));
        ++c.replays;
        return;
    }
    if (c.key != key) {
        if (c.exec) (void)hipGraphExecDestroy(c.exec);
        c.exec = nullptr;
        c.key = key;
        enqueue();
        return;
    }
    hipGraph_t graph = nullptr;
    ck(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
    try {
        enqueue();
    } catch (...) {
        (void)hipStreamEndCapture(st_, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
    }
    ck(hipStreamEndCapture(st_, &graph));
    const hipError_t e = hipGraphInstantiate(&c.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) c.exec = nullptr;
    ck(e);
    ++c.captures;
    ck(hipGraphLaunch(c.exec, st_));
    ++c.replays;
}

int Engine::shape_for(const std::vector<int>& key, bool listed) {
    if (force_shape >= 1 && force_shape <= shape_count(listed)) return force_shape - 1;
    if (force_shape > shape_count(listed)) return force_shape - 4;   // the same shape, row-major
    if (!autotune) return 0;
    const TuneState& t = tuned_[key];
    return t.chosen >= 0 ? t.chosen : t.next % shape_count(listed);
}

void Engine::shape_timed(const std::vector<int>& key, bool listed, int cand, float ms) {
    if (!autotune || force_shape != 0) return;
    TuneState& t = tuned_[key];
    if (t.chosen >= 0) return;
    t.samples[cand].push_back(ms);
    ++t.next;
    const int nc = shape_count(listed);
    TuneReport rep;
    rep.have = true;
    rep.settled = t.next >= nc * kTuneSamples;
    double best = 1e30;
    for (int i = 0; i < 6; ++i) {
        rep.ms[i] = -1.0;
        if (i >= nc || t.samples[i].empty()) continue;
        double sum = 0.0;
        for (float v : t.samples[i]) sum += v;
        rep.ms[i] = sum / (double)t.samples[i].size();
        if (rep.ms[i] < best) {
            best = rep.ms[i];
            rep.chosen = kShapes[i];
            if (rep.settled) t.chosen = i;
        }
    }
    last_tune = rep;
}

int Engine::set_subset_map(int gw, int gh, const std::vector<int>& dims) {
    if (dims.empty()) {
        subset_map_.clear();
        map_gw_ = map_gh_ = 0;
        return 0;
    }
    if (!size_map_ok(gw, gh, dims)) return kErrInit;
    subset_map_ = dims;
    map_gw_ = gw;
    map_gh_ = gh;
    return 0;
}

bool Engine::size_map_ok(int cols, int rows, const std::vector<int>& dims) {
    if (dims.empty()) return true;
    if (cols <= 0 || rows <= 0 || dims.size() != (size_t)cols * rows) return false;
    for (int d : dims)
        if (d < 3 || d % 2 == 0 || ppl_for(d) < 0) return false;
    return true;
}

int Engine::set_region_map(int rw, int rh, const std::vector<int>& dims) {
    if (!size_map_ok(rw, rh, dims)) return kErrInit;
    if (dims.empty()) {
        region_map_.clear();
        region_w_ = region_h_ = 0;
        return 0;
    }
    region_map_ = dims;
    region_w_ = rw;
    region_h_ = rh;
    return 0;
}

int Engine::ensure_grid(const Params& p, const uint8_t* mask, const std::vector<int>* dims, bool& cached) {
    const int gw = p.rect_w / p.step, gh = p.rect_h / p.step;
    const size_t n = (size_t)w_ * h_;
    const bool mask_same = (!mask && grid_mask_.empty()) ||
                           (mask && grid_mask_.size() == n && std::memcmp(grid_mask_.data(), mask, n) == 0);
    cached = grid_valid_ && grid_p_.rect_x == p.rect_x && grid_p_.rect_y == p.rect_y &&
             grid_p_.rect_w == p.rect_w && grid_p_.rect_h == p.rect_h && grid_p_.step == p.step &&
             grid_p_.subset == p.subset && mask_same && gw_ == gw && gh_ == gh && grid_prec_ == prec &&
             (dims ? grid_dims_ == *dims : grid_dims_.empty());
    if (cached) return 0;
    grid_valid_ = false;
    const int N = gw * gh;
    if (!skip_d_.reserve(N) || !pool_.reserve(N) || !counters_.reserve(32)) return kErrInit;
    const uint8_t* md = nullptr;
    if (mask) {
        if (!mask_d_.reserve(n)) return kErrInit;
        ck(hipMemcpyAsync(mask_d_.p, mask, n, hipMemcpyHostToDevice, st_));
        md = mask_d_.p;
    }
    const int32_t* dd = nullptr;
    classes_.clear();
    if (dims) {
        // per-node sizes, and the node lists of each size for Path A
        if (!dims_d_.reserve(N)) return kErrInit;
        ck(hipMemcpyAsync(dims_d_.p, dims->data(), sizeof(int32_t) * N, hipMemcpyHostToDevice, st_));
        dd = dims_d_.p;
        std::vector<int> sizes(dims->begin(), dims->end());
        std::sort(sizes.begin(), sizes.end());
        sizes.erase(std::unique(sizes.begin(), sizes.end()), sizes.end());
        std::vector<int32_t> lists;
        for (int d : sizes) {
            const int off = (int)lists.size();
            for (int i = 0; i < N; ++i)
                if ((*dims)[i] == d) lists.push_back(i);
            classes_.push_back(DimClass{d, off, (int)lists.size() - off});
        }
        if (!class_nodes_d_.reserve(lists.size())) return kErrInit;
        ck(hipMemcpyAsync(class_nodes_d_.p, lists.data(), sizeof(int32_t) * lists.size(), hipMemcpyHostToDevice,
                          st_));
    } else {
        classes_.push_back(DimClass{p.subset, -1, N});
    }
    ck(hipMemsetAsync(counters_.p, 0, sizeof(int32_t), st_));
    ck(launch_scan(md, w_, h_, p.rect_x, p.rect_y, p.step, gw, gh, p.subset, skip_d_.p, counters_.p, st_, dd));
    int total = 0;
    skip_h_.resize(N);
    ck(hipMemcpyAsync(&total, counters_.p, sizeof(int), hipMemcpyDeviceToHost, st_));
    ck(hipMemcpyAsync(skip_h_.data(), skip_d_.p, N, hipMemcpyDeviceToHost, st_));
    ck(hipStreamSynchronize(st_));
    total_valid_ = total;
    if (total == 0) return kErrRoi;
    PoolOpts po;
    po.mfma = prec.mfma_pool;
    po.dims = dd;
    if (prec.grad16) {
        const int fmt = prec.grad_bf16 ? 1 : 0;
        if (g16_fmt_ != fmt) {
            if (!gx16_.reserve(n) || !gy16_.reserve(n)) return kErrInit;
            if (fmt == 1) {
                ck(launch_to_bf16(gx_.p, n, gx16_.p, st_));
                ck(launch_to_bf16(gy_.p, n, gy16_.p, st_));
            } else {
                ck(launch_to_f16(gx_.p, n, gx16_.p, st_));
                ck(launch_to_f16(gy_.p, n, gy16_.p, st_));
            }
            g16_fmt_ = fmt;
        }
        po.gx16 = gx16_.p;
        po.gy16 = gy16_.p;
        po.bf16 = fmt == 1;
    }
    if (prec.chol64) {
        if (!chol_.reserve((size_t)N * kCholStride)) return kErrInit;
        po.chol = chol_.p;
    }
    if (!pool_raw_.reserve(pool_raw_bytes(N))) return kErrInit;
    ck(launch_pool(I_.p, gx_.p, gy_.p, w_, h_, skip_d_.p, gw, gh, p.rect_x, p.rect_y, p.step, p.subset, true,
                   pool_.p, pool_raw_.p, st_, po));
    ck(hipStreamSynchronize(st_));
    grid_prec_ = prec;
    grid_dims_ = dims ? *dims : std::vector<int>();
    grid_p_ = p;
    grid_mask_ = mask ? std::vector<uint8_t>(mask, mask + n) : std::vector<uint8_t>();
    gw_ = gw;
    gh_ = gh;
    grid_valid_ = true;
    return 0;
}

int Engine::load_seeder(const std::string& path) {
    if (!seeder_) seeder_ = std::make_unique<Seeder>();
    return seeder_->load(path);
}

void Engine::set_precision_args(SolveArgs& a) const {
    if (grid_prec_.grad16) {
        a.refGx16 = gx16_.p;
        a.refGy16 = gy16_.p;
        a.grad_bf16 = grid_prec_.grad_bf16 ? 1 : 0;
    }
    if (grid_prec_.chol64) a.chol = chol_.p;
}

void Engine::ensure_class_streams(size_t nc) {
    while (class_ev_.size() < 2 * nc + 2) {
        hipEvent_t e;
        ck(hipEventCreate(&e));
        class_ev_.push_back(e);
    }
    while (class_st_.size() + 1 < nc) {
        hipStream_t s2;
        ck(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
        class_st_.push_back(s2);
    }
    if (!class_nbail_.reserve(std::max<size_t>(nc, 1))) throw HipFail{hipErrorOutOfMemory};
}

int Engine::solve_tasks(const std::vector<Task>& tasks, std::vector<SolveOut>& out, bool use6, int dim, float tol) {
    if (grid_dims_.empty()) return solve_tasks_dim(tasks, out, use6, dim, tol);
    // mixed sizes: the tasks grouped by their node's size, one launch per
    // size, the groups after the first on streams of their own (forked from
    // the engine's stream and joined back), with one upload and one download
    out.assign(tasks.size(), SolveOut{});
    struct Seg {
        int dim, off, n;
    };
    std::vector<Seg> segs;
    std::vector<Task> sorted;
    std::vector<size_t> at;
    for (const DimClass& c : classes_) {
        const int off = (int)sorted.size();
        for (size_t k = 0; k < tasks.size(); ++k)
            if (tasks[k].pool >= 0 && (size_t)tasks[k].pool < grid_dims_.size() && grid_dims_[tasks[k].pool] == c.dim) {
                sorted.push_back(tasks[k]);
                at.push_back(k);
            }
        if ((int)sorted.size() > off) segs.push_back(Seg{c.dim, off, (int)sorted.size() - off});
    }
    if (segs.empty()) return 0;
    const int n = (int)sorted.size();
    const size_t ns = segs.size();
    if (!tasks_d_.reserve(n) || !out_d_.reserve(n) || !bail_.reserve(n) || !counters_.reserve(32)) return kErrInit;
    if (!pin_tasks_.reserve(n) || !pin_res_.reserve(n) || !pin_handed_.reserve(ns)) return kErrInit;
    ensure_class_streams(ns);
    std::memcpy(pin_tasks_.p, sorted.data(), sizeof(Task) * n);
    ck(hipMemcpyAsync(tasks_d_.p, pin_tasks_.p, sizeof(Task) * n, hipMemcpyHostToDevice, st_));
    hipEvent_t ev_fork = class_ev_[2 * ns];
    ck(hipEventRecord(ev_fork, st_));
    std::vector<SolveArgs> sargs(ns);
    for (size_t i = 0; i < ns; ++i) {
        hipStream_t cs = i == 0 ? st_ : class_st_[i - 1];
        if (i > 0) ck(hipStreamWaitEvent(cs, ev_fork, 0));
        SolveArgs a = {};
        a.tasks = tasks_d_.p + segs[i].off;
        a.out = out_d_.p + segs[i].off;
        a.nitems = segs[i].n;
        a.bail = bail_.p + segs[i].off;
        a.nbail = i == 0 ? counters_.p + 8 : class_nbail_.p + i;
        a.pool = pool_.p;
        a.refI = I_.p;
        a.refGx = gx_.p;
        a.refGy = gy_.p;
        set_precision_args(a);
        set_def_args(a);
        a.w = w_;
        a.h = h_;
        a.dim = segs[i].dim;
        a.lm = 1;
        a.tol = tol;
        sargs[i] = a;
        ck(launch_solve(a, use6, def_f32_ready_, cs, kSolveFast));
    }
    for (size_t i = 0; i < ns; ++i) {
        hipStream_t cs = i == 0 ? st_ : class_st_[i - 1];
        ck(launch_solve(sargs[i], use6, def_f32_ready_, cs, kSolveFull));
        if (i > 0) {
            ck(hipEventRecord(class_ev_[2 * i + 1], cs));
            ck(hipStreamWaitEvent(st_, class_ev_[2 * i + 1], 0));
        }
    }
    ck(hipMemcpyAsync(pin_res_.p, out_d_.p, sizeof(SolveOut) * n, hipMemcpyDeviceToHost, st_));
    ck(hipMemcpyAsync(pin_handed_.p, counters_.p + 8, sizeof(int32_t), hipMemcpyDeviceToHost, st_));
    if (ns > 1)
        ck(hipMemcpyAsync(pin_handed_.p + 1, class_nbail_.p + 1, sizeof(int32_t) * (ns - 1), hipMemcpyDeviceToHost,
                          st_));
    ck(hipStreamSynchronize(st_));
    for (int j = 0; j < n; ++j) out[at[j]] = pin_res_.p[j];
    for (size_t i = 0; i < ns; ++i) launch_log.emplace_back(segs[i].n, pin_handed_.p[i]);
    return 0;
}

int Engine::solve_tasks_dim(const std::vector<Task>& tasks, std::vector<SolveOut>& out, bool use6, int dim,
                            float tol) {
    const int n = (int)tasks.size();
    out.assign(n, SolveOut{});
    if (n == 0) return 0;
    if (!tasks_d_.reserve(n) || !out_d_.reserve(n) || !bail_.reserve(n) || !counters_.reserve(32)) return kErrInit;
    // through page-locked staging: the copies are DMA transfers the stream
    // orders, with no staging by the runtime
    if (!pin_tasks_.reserve(n) || !pin_res_.reserve(n) || !pin_handed_.reserve(1)) return kErrInit;
    std::memcpy(pin_tasks_.p, tasks.data(), sizeof(Task) * n);
    ck(hipMemcpyAsync(tasks_d_.p, pin_tasks_.p, sizeof(Task) * n, hipMemcpyHostToDevice, st_));
    SolveArgs a = {};
    a.tasks = tasks_d_.p;
    a.out = out_d_.p;
    a.nitems = n;
    a.bail = bail_.p;
    a.nbail = counters_.p + 8;
    a.pool = pool_.p;
    a.refI = I_.p;
    a.refGx = gx_.p;
    a.refGy = gy_.p;
    set_precision_args(a);
    set_def_args(a);
    a.w = w_;
    a.h = h_;
    a.dim = dim;
    a.lm = 1;
    a.tol = tol;
    ck(launch_solve(a, use6, def_f32_ready_, st_));
    ck(hipMemcpyAsync(pin_res_.p, out_d_.p, sizeof(SolveOut) * n, hipMemcpyDeviceToHost, st_));
    ck(hipMemcpyAsync(pin_handed_.p, a.nbail, sizeof(int32_t), hipMemcpyDeviceToHost, st_));
    ck(hipStreamSynchronize(st_));
    std::memcpy(out.data(), pin_res_.p, sizeof(SolveOut) * n);
    launch_log.emplace_back(n, pin_handed_.p[0]);
    return 0;
}

int Engine::solve_tasks_spec(const std::vector<Task>& t1, const std::vector<Task>& t2, std::vector<SolveOut>& o1,
                             std::vector<SolveOut>& o2, std::vector<uint8_t>& handed, bool use6, int dim,
                             float tol) {
    const int n = (int)t1.size();
    o1.assign(n, SolveOut{});
    o2.assign(n, SolveOut{});
    handed.assign(n, 0);
    if (n == 0) return 0;
    if ((int)t2.size() != n) return kErrInit;
    const size_t n2 = 2 * (size_t)n;
    if (!tasks_d_.reserve(n2) || !out_d_.reserve(n2) || !bail_.reserve(n2) || !counters_.reserve(32)) return kErrInit;
    if (!pin_tasks_.reserve(n2) || !pin_res_.reserve(n2) || !pin_handed_.reserve(2 + (size_t)n)) return kErrInit;
    ensure_class_streams(2);
    std::memcpy(pin_tasks_.p, t1.data(), sizeof(Task) * n);
    std::memcpy(pin_tasks_.p + n, t2.data(), sizeof(Task) * n);
    ck(hipMemcpyAsync(tasks_d_.p, pin_tasks_.p, sizeof(Task) * n2, hipMemcpyHostToDevice, st_));
    SolveArgs a = {};
    a.tasks = tasks_d_.p;
    a.out = out_d_.p;
    a.nitems = n;
    a.bail = bail_.p;
    a.nbail = counters_.p + 8;
    a.pool = pool_.p;
    a.refI = I_.p;
    a.refGx = gx_.p;
    a.refGy = gy_.p;
    set_precision_args(a);
    set_def_args(a);
    a.w = w_;
    a.h = h_;
    a.dim = dim;
    a.lm = 1;
    a.tol = tol;
    ck(launch_solve(a, use6, def_f32_ready_, st_, kSolveFast));
    ck(hipMemcpyAsync(pin_handed_.p, a.nbail, sizeof(int32_t), hipMemcpyDeviceToHost, st_));
    ck(hipStreamSynchronize(st_));
    const int nb = pin_handed_.p[0];
    ck(launch_solve(a, use6, def_f32_ready_, st_, kSolveFull));
    if (nb > 0) {
        // the second set over the handed-on items: item id k is task k of the
        // second set (tasks_d_ + n) and result slot k of its half of out_d_;
        // the first set's fast path is complete (the stream was synchronised)
        hipStream_t cs = class_st_[0];
        SolveArgs b = a;
        b.tasks = tasks_d_.p + n;
        b.out = out_d_.p + n;
        b.items = bail_.p;
        b.nitems = nb;
        b.bail = bail_.p + n;
        b.nbail = class_nbail_.p + 1;
        ck(launch_solve(b, use6, def_f32_ready_, cs, kSolveBoth));
        ck(hipEventRecord(class_ev_[1], cs));
        ck(hipStreamWaitEvent(st_, class_ev_[1], 0));
        ck(hipMemcpyAsync(pin_handed_.p + 1, class_nbail_.p + 1, sizeof(int32_t), hipMemcpyDeviceToHost, st_));
        ck(hipMemcpyAsync(pin_handed_.p + 2, bail_.p, sizeof(int32_t) * nb, hipMemcpyDeviceToHost, st_));
    }
    ck(hipMemcpyAsync(pin_res_.p, out_d_.p, sizeof(SolveOut) * n2, hipMemcpyDeviceToHost, st_));
    ck(hipStreamSynchronize(st_));
    std::memcpy(o1.data(), pin_res_.p, sizeof(SolveOut) * n);
    launch_log.emplace_back(n, nb);
    if (nb > 0) {
        for (int k = 0; k < nb; ++k) {
            const int id = pin_handed_.p[2 + k];
            if (id < 0 || id >= n) return kErrInit;
            handed[id] = 1;
            o2[id] = pin_res_.p[n + id];
        }
        launch_log.emplace_back(nb, pin_handed_.p[1]);
    }
    return 0;
}

int Engine::solve_tasks_device(const Task* tasks, int n, SolveOut* out, bool use6, int dim, float tol) {
    if (n <= 0) return 0;
    if (!bail_.reserve(n) || !counters_.reserve(32)) return kErrInit;
    SolveArgs a = {};
    a.tasks = tasks;
    a.out = out;
    a.nitems = n;
    a.bail = bail_.p;
    a.nbail = counters_.p + 8;
    a.pool = pool_.p;
    a.refI = I_.p;
    a.refGx = gx_.p;
    a.refGy = gy_.p;
    set_precision_args(a);
    set_def_args(a);
    a.w = w_;
    a.h = h_;
    a.dim = dim;
    a.lm = 1;
    a.tol = tol;
    ck(launch_solve(a, use6, def_f32_ready_, st_));
    return 0;
}

FloodDev Engine::flood_dev(int gw, int gh) {
    const size_t NN = (size_t)gw * gh;
    FloodDev f;
    if (!fl_u8_.reserve(4 * NN) || !fl_w_.reserve(7 * NN) || !fl_int_.reserve(9 * NN) || !fl_key_.reserve(3 * NN) ||
        !fl_tasks_.reserve(NN) || !fl_res_.reserve(3 * NN) || !fl_counters_.reserve(kFcCount) ||
        !pin_fl_.reserve(kFcCount + 1))
        throw HipFail{hipErrorOutOfMemory};
    const size_t temp = flood_sort_temp_bytes((int)NN);
    if (!fl_sort_temp_.reserve(temp)) throw HipFail{hipErrorOutOfMemory};
    f.gw = gw;
    f.gh = gh;
    f.ok = fl_u8_.p;
    f.fin = fl_u8_.p + NN;
    f.committed = fl_u8_.p + 2 * NN;
    f.tried = (int8_t*)(fl_u8_.p + 3 * NN);
    f.w = fl_w_.p;
    int* ip = fl_int_.p;
    f.stamp = ip;
    f.dist = ip + NN;
    f.cand = ip + 2 * NN;
    f.sources = ip + 3 * NN;
    f.level_a = ip + 4 * NN;
    f.level_b = ip + 5 * NN;
    f.reached = ip + 6 * NN;
    f.tdist = ip + 7 * NN;
    f.commit_nodes = ip + 8 * NN;
    f.key = fl_key_.p;
    f.skeys = fl_key_.p + NN;
    f.skeys_sorted = fl_key_.p + 2 * NN;
    f.tasks = fl_tasks_.p;
    f.commit_res = fl_res_.p + NN;
    f.commit_list = fl_res_.p + 2 * NN;
    f.counters = fl_counters_.p;
    return f;
}

// match_anchor (src/pipeline/incremental_math.cpp:237-272) for a batch of
// stored subsets. The patches are laid side by side on a canvas as wide as the
// frame, so the subset solver reads the canvas as its reference plane (the
// subset pixels lie at least the anchor margin inside their patch, so their
// gradients are the patch's own) and the last run's frame as its deformed
// plane; a subset centred at (cx, cy) on the canvas starts at the node's guess
// less its offset from the patch centre.
int Engine::cut_patches(const std::vector<int32_t>& entries, int side, uint8_t* store) {
    const int n = (int)(entries.size() / 3);
    if (n <= 0) return 0;
    if (!def_u8_.p || !store || side <= 0) return kErrInit;
    try {
        if (!patch_list_.reserve(entries.size())) return kErrInit;
        ck(hipMemcpyAsync(patch_list_.p, entries.data(), sizeof(int32_t) * entries.size(), hipMemcpyHostToDevice, st_));
        ck(launch_cut_patches(def_u8_.p, w_, patch_list_.p, n, side, store, st_));
        // the entries' host buffer may go before the copy would otherwise run
        ck(hipStreamSynchronize(st_));
        return 0;
    } catch (const HipFail&) {
        return kErrInit;
    }
}

int Engine::match_anchors(std::vector<AnchorMatch>& reqs, const uint8_t* store, int subset, bool use6, float tol) {
    for (auto& r : reqs) r.ok = false;
    if (reqs.empty()) return 0;
    if (w_ <= 0 || h_ <= 0 || !def_u8_.p || !store) return kErrInit;
    const int P = reqs[0].side;
    for (const auto& r : reqs)
        if (r.side != P || r.slot < 0) return kErrInit;
    const int cols = w_ / P, rows_max = h_ / P;
    if (cols < 1 || rows_max < 1 || ppl_for(subset) < 0) return 0;   // no patch fits the frame: nothing matches
    const int per_batch = cols * rows_max;
    try {
        for (size_t first = 0; first < reqs.size(); first += per_batch) {
            const int n = (int)std::min<size_t>(per_batch, reqs.size() - first);
            const int rows = (n + cols - 1) / cols, Hc = rows * P;
            const size_t npx = (size_t)w_ * Hc;
            std::vector<int32_t> cxy(2 * (size_t)n), slots(n);
            std::vector<Task> tasks(n);
            for (int k = 0; k < n; ++k) {
                const AnchorMatch& r = reqs[first + k];
                const int ox = (k % cols) * P, oy = (k / cols) * P;
                slots[k] = r.slot;
                Task t = {};
                t.cx = ox + P / 2;
                t.cy = oy + P / 2;
                t.pool = k;
                t.mode = NO_SEARCH;
                t.g[0] = r.guess_x - r.dx - (float)t.cx;
                t.g[1] = r.guess_y - r.dy - (float)t.cy;
                t.node = k;
                tasks[k] = t;
                cxy[2 * k] = t.cx;
                cxy[2 * k + 1] = t.cy;
            }
            DevBuf<uint8_t> dcan;
            DevBuf<float> planes;
            DevBuf<int32_t> dcxy;
            DevBuf<PoolEntry> pool;
            if (!dcan.reserve(npx) || !planes.reserve(3 * npx) || !dcxy.reserve(2 * (size_t)n) || !pool.reserve(n) ||
                !tasks_d_.reserve(n) || !out_d_.reserve(n) || !bail_.reserve(n) || !counters_.reserve(32))
                return kErrInit;
            float* cI = planes.p;
            float* cgx = cI + npx;
            float* cgy = cgx + npx;
            if (!patch_list_.reserve((size_t)n)) return kErrInit;
            ck(hipMemcpyAsync(patch_list_.p, slots.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, st_));
            ck(launch_place_patches(store, P, patch_list_.p, n, cols, w_, rows, dcan.p, st_));
            ck(hipMemcpyAsync(dcxy.p, cxy.data(), sizeof(int32_t) * 2 * n, hipMemcpyHostToDevice, st_));
            ck(hipMemcpyAsync(tasks_d_.p, tasks.data(), sizeof(Task) * n, hipMemcpyHostToDevice, st_));
            ck(launch_prep_ref_u8(dcan.p, nullptr, w_, Hc, cI, cgx, cgy, st_));
            ck(launch_pool_list(cI, cgx, cgy, w_, Hc, dcxy.p, n, subset, true, pool.p, st_));
            SolveArgs a = {};
            a.tasks = tasks_d_.p;
            a.out = out_d_.p;
            a.nitems = n;
            a.bail = bail_.p;
            a.nbail = counters_.p + 8;
            a.pool = pool.p;
            a.refI = cI;
            a.refGx = cgx;
            a.refGy = cgy;
            set_def_args(a);
            a.w = w_;
            a.h = h_;
            a.dim = subset;
            a.lm = 1;
            a.tol = tol;
            ck(launch_solve(a, use6, def_f32_ready_, st_));
            std::vector<SolveOut> res(n);
            ck(hipMemcpyAsync(res.data(), out_d_.p, sizeof(SolveOut) * n, hipMemcpyDeviceToHost, st_));
            ck(hipStreamSynchronize(st_));
            for (int k = 0; k < n; ++k) {
                AnchorMatch& r = reqs[first + k];
                const SolveOut& o = res[k];
                if (!o.init || o.status != 0 || !(o.corr <= kCorrAccept)) continue;
                // the node, (dx, dy) from the patch centre, through the matched map
                const float u = o.p[0] + (float)(tasks[k].cx - r.ax), v = o.p[1] + (float)(tasks[k].cy - r.ay);
                const float nx = (float)r.ax + u + (1.0f + o.p[2]) * r.dx + o.p[3] * r.dy;
                const float ny = (float)r.ay + v + o.p[4] * r.dx + (1.0f + o.p[5]) * r.dy;
                if (!(std::hypot(nx - r.guess_x, ny - r.guess_y) <= r.max_shift)) continue;
                r.ok = true;
                r.x = nx;
                r.y = ny;
                r.corr = o.corr;
            }
        }
        return 0;
    } catch (const HipFail&) {
        return kErrInit;
    }
}

int Engine::run(const uint8_t* def, bool on_device, const uint8_t* mask, const Params& pin, float* out,
                int out_capacity, float* metrics, int metrics_len, Cancel& cancel,
                const std::function<void(int)>& progress, float tol) {
    const auto t_total = clk::now();
    tm_ = Timings{};
    launch_log.clear();
    if (!out || !have_ref_ || !def) return kErrInit;
    Params p = pin;
    const int rc0 = validate(p, w_, h_);
    if (rc0 != 0) return rc0;
    if (ppl_for(p.subset) < 0) return kErrInit;   // larger than the register-resident solver takes
    if (out_capacity < 0) out_capacity = 0;
    try {
        const size_t npx = (size_t)w_ * h_;
        const auto t_prep = clk::now();
        if (!def_u8_.reserve(npx)) return kErrInit;
        ck(hipMemcpyAsync(def_u8_.p, def, npx, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st_));
        def_f32_ready_ = false;
        if (def_f32) {
            if (!def_f32_.reserve(npx)) return kErrInit;
            ck(launch_u8_to_f32(def_u8_.p, npx, def_f32_.p, st_));
            def_f32_ready_ = true;
        }
        def_coef_ready_ = false;
        if (def_coef && def_f32_ready_ && npx < kCoefMaxPixels) {
            if (!def_coef_.reserve(npx)) return kErrInit;
            ck(launch_row_cubic(def_u8_.p, w_, h_, def_coef_.p, st_));
            def_coef_ready_ = true;
        }
        tm_.prep = ms_since(t_prep);
        const int gw = p.rect_w / p.step, gh = p.rect_h / p.step, N = gw * gh;
        if (vsg_count(p.step, p.strain_window) < 3) return kErrStrainWindow;
        if (cancel.requested()) return kErrCancelled;

        const auto t_pre = clk::now();
        bool cached = false;
        const bool mixed = !subset_map_.empty() && map_gw_ == gw && map_gh_ == gh;
        const std::vector<int>* dims = mixed ? &subset_map_ : nullptr;
        if (!dims && !region_map_.empty()) {
            region_dims_.resize((size_t)N);
            for (int gy = 0; gy < gh; ++gy)
                for (int gx = 0; gx < gw; ++gx) {
                    const int64_t x = p.rect_x + (int64_t)gx * p.step, y = p.rect_y + (int64_t)gy * p.step;
                    const int rx = (int)std::min<int64_t>(x * region_w_ / w_, region_w_ - 1);
                    const int ry = (int)std::min<int64_t>(y * region_h_ / h_, region_h_ - 1);
                    region_dims_[(size_t)gy * gw + gx] = region_map_[(size_t)ry * region_w_ + rx];
                }
            dims = &region_dims_;
        }
        const int rg = ensure_grid(p, mask, dims, cached);
        if (rg != 0) return rg;
        tm_.prepass = cached ? 0.0 : ms_since(t_pre);
        const int total_valid = total_valid_;
        if (cancel.requested()) return kErrCancelled;

        Ticker ticker(progress, total_valid);
        const size_t NN = (size_t)N;
        if (!gplanes_.reserve(7 * NN) || !gflags_.reserve(7 * NN) || !giters_.reserve(NN) || !guess_.reserve(6 * NN) ||
            !tmp_.reserve(6 * NN) || !strain_.reserve(3 * NN) || !bail_.reserve(NN) || !counters_.reserve(32))
            return kErrInit;
        GridDev g;
        g.u = gplanes_.p;
        g.v = g.u + NN;
        g.ux = g.v + NN;
        g.uy = g.ux + NN;
        g.vx = g.uy + NN;
        g.vy = g.vx + NN;
        g.corr = g.vy + NN;
        g.solved = gflags_.p;
        g.tag = g.solved + NN;
        g.simplex = g.tag + NN;
        uint8_t* in_mesh = g.simplex + NN;
        uint8_t* valid_d = in_mesh + NN;
        uint8_t* ok_a = valid_d + NN;
        uint8_t* ok_b = ok_a + NN;
        (void)ok_b;
        g.iters = giters_.p;
        ck(launch_grid_init(g, skip_d_.p, N, st_));
        std::vector<uint8_t> solved_h(skip_h_);   // host mirror of the solved flags

        // ---- seeding: global model + anchor lattice -------------------------
        const auto t_seed = clk::now();
        const int L = std::min(pyramid_levels(p.rect_w, p.rect_h), pyr_ref_levels_);
        // the learned seeder reads level 2 of both pyramids
        const bool cnn_ok = learned_seed && seeder_ready() && pyr_ref_levels_ >= SeederShape::kLevel;
        const int Lb = cnn_ok ? std::max(L, (int)SeederShape::kLevel) : L;
        seed_stats = SeedStats{};
        PyrDev pd = {};
        pd.r0 = ref_u8_.p;
        pd.d0 = def_u8_.p;
        pd.w[0] = w_;
        pd.h[0] = h_;
        int w2 = 0, h2 = 0;
        {
            int pw = w_, ph = h_;
            for (int l = 1; l <= Lb; ++l) {
                const int nw = pw / 2, nh = ph / 2;
                if (!pyr_d_[l].reserve((size_t)nw * nh)) return kErrInit;
                if (l == 1) ck(launch_downsample_u8(def_u8_.p, pw, ph, pyr_d_[1].p, st_));
                else ck(launch_downsample_f32(pyr_d_[l - 1].p, pw, ph, pyr_d_[l].p, st_));
                pd.r[l] = pyr_r_[l].p;
                pd.d[l] = pyr_d_[l].p;
                pd.w[l] = nw;
                pd.h[l] = nh;
                pw = nw;
                ph = nh;
            }
            if (cnn_ok) {
                w2 = pd.w[SeederShape::kLevel];
                h2 = pd.h[SeederShape::kLevel];
            }
        }
        pd.levels = L;
        std::vector<int> centres;
        patch_centres(p.rect_x, p.rect_y, p.rect_w, p.rect_h, centres);
        const int npatch = (int)centres.size() / 2;
        if (!centres_d_.reserve(centres.size()) || !patch_d_.reserve(npatch) ||
            !patch_scratch_.reserve(patch_scratch_bytes(npatch)))
            return kErrInit;
        ck(hipMemcpyAsync(centres_d_.p, centres.data(), sizeof(int) * centres.size(), hipMemcpyHostToDevice, st_));
        std::vector<PatchOut> po(npatch);
        GlobalModel gm;
        int pass = 0;
        for (const int rad : {kCoarseRadius, kCoarseRadiusWide}) {
            if (!global_search) break;
            pd.coarse_radius = rad;
            ck(launch_patch_search(pd, centres_d_.p, npatch, patch_d_.p, patch_scratch_.p, st_));
            ck(hipMemcpyAsync(po.data(), patch_d_.p, sizeof(PatchOut) * npatch, hipMemcpyDeviceToHost, st_));
            ck(hipStreamSynchronize(st_));
            gm = fit_global_model(centres, po, p.rect_x, p.rect_y, p.rect_w, p.rect_h);
            ++pass;
            // the translation search's untouched templates decorrelate long
            // before a quarter turn: a model it fits with a larger turn is a
            // false lock of mismatched patches
            if (gm.locked && std::fabs(model_turn_deg(gm)) > kTranslationMaxTurnDeg) gm = GlobalModel{};
            if (gm.locked) break;
        }
        // Motions the translation passes cannot follow (a large rotation,
        // rotation with scale): rotation-invariant matching of 5 x 5 patch
        // centres over the whole deformed image at the top level, a rough
        // affine through their consensus, then the translation search again
        // through that affine's linear map, started at its predicted shifts.
        if (global_search && !gm.locked && pd.w[L] > 2 * kCanonMargin + 2 && pd.h[L] > 2 * kCanonMargin + 2) {
            const int g5 = 5;
            std::vector<int> cl;
            for (int j = 0; j < g5; ++j)
                for (int i = 0; i < g5; ++i) {
                    cl.push_back((p.rect_x + (int)std::lround((i + 0.5) * p.rect_w / g5)) >> L);
                    cl.push_back((p.rect_y + (int)std::lround((j + 0.5) * p.rect_h / g5)) >> L);
                }
            const int nc = g5 * g5;
            const size_t nscores = canon_scores_floats(pd.w[L], pd.h[L], nc);
            if (!canon_centres_d_.reserve(cl.size()) || !canon_desc_.reserve((size_t)nc * kCanonSamples + nc) ||
                !canon_scores_.reserve(nscores) || !canon_out_.reserve(nc) || !canon_start_.reserve(2 * (size_t)npatch))
                return kErrInit;
            ck(hipMemcpyAsync(canon_centres_d_.p, cl.data(), sizeof(int) * cl.size(), hipMemcpyHostToDevice, st_));
            ck(launch_canon_match(pd, L, canon_centres_d_.p, nc, canon_desc_.p, canon_desc_.p + (size_t)nc * kCanonSamples,
                                  canon_scores_.p, canon_out_.p, st_));
            std::vector<CanonOut> co(nc);
            ck(hipMemcpyAsync(co.data(), canon_out_.p, sizeof(CanonOut) * nc, hipMemcpyDeviceToHost, st_));
            ck(hipStreamSynchronize(st_));
            const double mcx = p.rect_x + 0.5 * p.rect_w, mcy = p.rect_y + 0.5 * p.rect_h;
            double ra[6];
            if (canon_affine(cl, co, L, mcx, mcy, ra)) {
                std::vector<int> start(2 * (size_t)npatch);
                const double sc = (double)(1 << L);
                for (int k = 0; k < npatch; ++k) {
                    const double dx = centres[2 * k] - mcx, dy = centres[2 * k + 1] - mcy;
                    start[2 * k] = (int)std::lround((ra[0] + ra[1] * dx + ra[2] * dy) / sc);
                    start[2 * k + 1] = (int)std::lround((ra[3] + ra[4] * dx + ra[5] * dy) / sc);
                }
                ck(hipMemcpyAsync(canon_start_.p, start.data(), sizeof(int) * start.size(), hipMemcpyHostToDevice,
                                  st_));
                PyrDev pw = pd;
                pw.warp = 1;
                pw.f[0] = 1.0 + ra[1];
                pw.f[1] = ra[2];
                pw.f[2] = ra[4];
                pw.f[3] = 1.0 + ra[5];
                pw.start = canon_start_.p;
                pw.coarse_radius = 2;
                ck(launch_patch_search(pw, centres_d_.p, npatch, patch_d_.p, patch_scratch_.p, st_));
                ck(hipMemcpyAsync(po.data(), patch_d_.p, sizeof(PatchOut) * npatch, hipMemcpyDeviceToHost, st_));
                ck(hipStreamSynchronize(st_));
                gm = fit_global_model(centres, po, p.rect_x, p.rect_y, p.rect_w, p.rect_h);
                pass = 3;
            }
        }
        {
            const double v[12] = {gm.locked ? 1.0 : 0.0, gm.full ? 1.0 : 0.0, gm.a[0], gm.a[1], gm.a[2], gm.a[3],
                                  gm.a[4], gm.a[5], gm.cx, gm.cy, (double)gm.patches, gm.locked ? (double)pass : 0.0};
            std::copy(v, v + 12, last_model);
        }

        // Learned seeding (no reference counterpart) runs the CNN lattice
        // when the global model finds a large motion (an ROI corner moving
        // more than 4 px beyond the centre's translation, or no lock). A
        // near-translation is the reference's plain lattice's case: it seeds
        // it first, and the CNN lattice follows only if that lattice's
        // anchors do not converge.
        bool use_cnn = cnn_ok && (cnn_always || !(gm.locked && !gm.full));
        // lattice (solve_anchor_seeds, full_field_anchors.cpp:283-574)
        int sx = 1, sy = 1, LW = 0, LH = 0;
        std::vector<int> lx, ly;
        auto plan = [&]() {
            // the learned seeder's lattice is 16 times denser (anchors 4 times
            // closer): its guesses cost one CNN pass each instead of a coarse
            // search, and the 8-neighbour median test (+-1 px on residuals)
            // holds where a field curves between anchors of the plain spacing
            plan_lattice(gw, gh, sx, sy, use_cnn ? 16 * kAnchorTarget : kAnchorTarget);
            lx.clear();
            ly.clear();
            for (int x = 0; x < gw; x += sx) lx.push_back(x);
            if (lx.back() != gw - 1) lx.push_back(gw - 1);
            for (int y = 0; y < gh; y += sy) ly.push_back(y);
            if (ly.back() != gh - 1) ly.push_back(gh - 1);
            LW = (int)lx.size();
            LH = (int)ly.size();
        };
        plan();
        struct Seeds {
            std::vector<AnchorRec> an;
            int quality = 0, attempted = 0, accepted = 0, solved = 0;
            float gU = 0, gV = 0, coverage = 0;
            bool model_rejected = false;
            std::vector<float> ref_x, ref_y, def_x, def_y;
        } seeds;
        Counters st_anchor, st_a, st_b;
        double icgn_ms_est = 0;
        seeds.gU = (float)gm.t_u;
        seeds.gV = (float)gm.t_v;
        auto solve_lattice = [&](bool sim, Seeds& r) -> float {
            r = Seeds{};
            r.gU = (float)gm.t_u;
            r.gV = (float)gm.t_v;
            const int na = LW * LH;
            r.an.assign(na, AnchorRec{});
            std::vector<Task> tasks;
            std::vector<int> which;
            for (int idx = 0; idx < na; ++idx) {
                const int gx = lx[idx % LW], gy = ly[idx / LW];
                if (solved_h[(size_t)gy * gw + gx]) continue;
                const int rx = p.rect_x + gx * p.step, ry = p.rect_y + gy * p.step;
                Task t = {};
                t.cx = rx;
                t.cy = ry;
                t.pool = gy * gw + gx;
                t.mode = gm.locked ? NO_SIMPLEX : AUTO_SEARCH;
                float gs[6] = {r.gU, r.gV, 0, 0, 0, 0};
                if (sim) gm.at(rx, ry, gs);
                std::copy(gs, gs + 6, t.g);
                t.node = t.pool;
                tasks.push_back(t);
                which.push_back(idx);
            }
            r.attempted = (int)tasks.size();
            std::vector<SolveOut> res;
            const auto t0 = clk::now();
            std::vector<Task> plain;
            if (use_cnn && !tasks.empty()) {
                // the CNN's warp of each anchor's patch pair, the deformed patch
                // placed by the global model's shift at the patch centre
                plain = tasks;
                const int n = (int)tasks.size(), S2 = SeederShape::kScale;
                std::vector<int> c2(2 * n), off(2 * n);
                std::vector<float> y(6 * (size_t)n);
                std::vector<double> px(n), py(n);
                for (int k = 0; k < n; ++k) {
                    c2[2 * k] = (int)std::lround((tasks[k].cx + 0.5) / S2);
                    c2[2 * k + 1] = (int)std::lround((tasks[k].cy + 0.5) / S2);
                    px[k] = (double)S2 * c2[2 * k] - 0.5;
                    py[k] = (double)S2 * c2[2 * k + 1] - 0.5;
                    float bg[6] = {0, 0, 0, 0, 0, 0};
                    if (gm.locked) gm.at(px[k], py[k], bg);
                    off[2 * k] = (int)std::lround(bg[0] / S2);
                    off[2 * k + 1] = (int)std::lround(bg[1] / S2);
                }
                seeder_->use_mfma = seeder_mfma;
                if (seeder_->infer(pyr_r_[SeederShape::kLevel].p, pyr_d_[SeederShape::kLevel].p, w2, h2, c2.data(),
                                   off.data(), n, y.data(), st_) != 0)
                    throw HipFail{hipErrorLaunchFailure};
                for (int k = 0; k < n; ++k) {
                    const float* q = &y[6 * (size_t)k];
                    const double U = (double)S2 * off[2 * k] + q[0], V = (double)S2 * off[2 * k + 1] + q[1];
                    const double dx = tasks[k].cx - px[k], dy = tasks[k].cy - py[k];
                    tasks[k].g[0] = (float)(U + q[2] * dx + q[3] * dy);
                    tasks[k].g[1] = (float)(V + q[4] * dx + q[5] * dy);
                    for (int j = 2; j < 6; ++j) tasks[k].g[j] = q[j];
                    // IC-GN from the CNN's warp; where it fails, the plain
                    // anchor solve below (the Simplex rescue first, if asked)
                    tasks[k].mode = cnn_rescue ? NO_SEARCH : NO_SIMPLEX;
                }
            }
            // Under the CNN guesses on one subset size, the plain solves of the
            // anchors whose CNN-seeded IC-GN the fast path handed on run while
            // the full solver rescues them (its slowest anchor sets its
            // launch's length); each anchor's result is the one the plain
            // solve after the rescue would give.
            std::vector<SolveOut> pre;
            std::vector<uint8_t> handed;
            if (!plain.empty() && grid_dims_.empty()) {
                if (solve_tasks_spec(tasks, plain, res, pre, handed, p.use6, p.subset, tol) != 0)
                    throw HipFail{hipErrorOutOfMemory};
            } else if (solve_tasks(tasks, res, p.use6, p.subset, tol) != 0) {
                throw HipFail{hipErrorOutOfMemory};
            }
            if (!plain.empty()) {
                // the plain anchor solve wherever the refined CNN guess failed
                std::vector<Task> retry;
                std::vector<int> ridx;
                int failed = 0;
                for (size_t k = 0; k < tasks.size(); ++k) {
                    const SolveOut& o = res[k];
                    if (o.init && o.status == 0 && o.corr <= kAnchorAcceptScore) continue;
                    ++failed;
                    if (!handed.empty() && handed[k]) {
                        res[k] = pre[k];
                        tasks[k] = plain[k];
                        continue;
                    }
                    retry.push_back(plain[k]);
                    ridx.push_back((int)k);
                }
                seed_stats.cnn += (int)tasks.size();
                seed_stats.cnn_ok += (int)tasks.size() - failed;
                seed_stats.fallback += failed;
                if (!retry.empty()) {
                    std::vector<SolveOut> rr;
                    if (solve_tasks(retry, rr, p.use6, p.subset, tol) != 0) throw HipFail{hipErrorOutOfMemory};
                    for (size_t i = 0; i < retry.size(); ++i) {
                        res[ridx[i]] = rr[i];
                        tasks[ridx[i]] = retry[i];
                    }
                }
            }
            icgn_ms_est += ms_since(t0);
            for (size_t k = 0; k < tasks.size(); ++k) {
                const SolveOut& o = res[k];
                if (!o.init) continue;
                st_anchor.record(o);
                if (o.status != 0) continue;
                AnchorRec& a = r.an[which[k]];
                a.rx = (float)tasks[k].cx;
                a.ry = (float)tasks[k].cy;
                std::copy(o.p, o.p + 6, a.p);
                a.corr = o.corr;
                a.gx = lx[which[k] % LW];
                a.gy = ly[which[k] / LW];
                a.node = tasks[k].pool;
                a.iters = o.iters;
                a.simplex = o.simplex != 0;
                a.vertex = o.corr <= kAnchorAcceptScore;
                a.output = o.corr <= kCorrAccept;
            }
            std::vector<float> ru(na, 0.0f), rv(na, 0.0f);
            for (int i = 0; i < na; ++i) {
                const AnchorRec& a = r.an[i];
                if (!a.vertex) continue;
                float gs[6] = {0, 0, 0, 0, 0, 0};
                if (sim) gm.at(a.rx, a.ry, gs);
                ru[i] = a.p[0] - gs[0];
                rv[i] = a.p[1] - gs[1];
            }
            mark_kept(r.an, LW, LH, ru, rv);
            // learned seeding (no reference counterpart): its lattice is dense
            // enough that a smooth field's variation between neighbours exceeds
            // the median test's 1 px, so an anchor that passed the output gate
            // -- a converged ZNSSD match -- is kept without it; looser vertices
            // still take the test, and the finished field takes its own
            if (use_cnn)
                for (AnchorRec& a : r.an)
                    if (a.output) a.kept = true;
            std::vector<float> us, vs;
            std::vector<std::pair<double, double>> pts;
            for (const AnchorRec& a : r.an) {
                if (!a.kept) continue;
                r.ref_x.push_back(a.rx);
                r.ref_y.push_back(a.ry);
                r.def_x.push_back(a.rx + a.p[0]);
                r.def_y.push_back(a.ry + a.p[1]);
                us.push_back(a.p[0]);
                vs.push_back(a.p[1]);
                pts.emplace_back(a.rx, a.ry);
                if (a.output) r.solved++;
            }
            r.accepted = (int)r.ref_x.size();
            if (r.accepted >= 3) {
                r.gU = median_of(us.data(), us.size());
                r.gV = median_of(vs.data(), vs.size());
                const float area = (float)(p.rect_w * p.rect_h);
                r.coverage = area > 0.0f ? (float)hull_area(pts) / area : 0.0f;
            }
            const float frac = r.attempted > 0 ? (float)r.accepted / (float)r.attempted : 0.0f;
            if (r.accepted >= kAnchorFullMinVertices && frac >= kAnchorFullFraction &&
                r.coverage >= kAnchorFullCoverage)
                r.quality = 2;
            else if (r.accepted >= 3 && r.coverage >= kAnchorSparseCoverage)
                r.quality = 1;
            return frac;
        };
        bool have_affine = false;
        double aff[6] = {0, 0, 0, 0, 0, 0}, acx = 0, acy = 0;
        auto seed_lattice = [&]() {
            have_affine = false;
            if (LW < 2 || LH < 2) return;
            if (!gm.full) {
                solve_lattice(false, seeds);
            } else {
                const float kept = solve_lattice(true, seeds);
                if (seeds.quality != 2 || kept < kSimilarityKeepFraction) {
                    solve_lattice(false, seeds);
                    seeds.model_rejected = true;
                }
            }
            bool refit = false;
            if (seeds.accepted >= kAnchorRefitMinPoints) {
                std::vector<double> xr, yr, ud, vd;
                for (size_t i = 0; i < seeds.ref_x.size(); ++i) {
                    xr.push_back(seeds.ref_x[i]);
                    yr.push_back(seeds.ref_y[i]);
                    ud.push_back((double)seeds.def_x[i] - seeds.ref_x[i]);
                    vd.push_back((double)seeds.def_y[i] - seeds.ref_y[i]);
                }
                acx = p.rect_x + 0.5 * p.rect_w;
                acy = p.rect_y + 0.5 * p.rect_h;
                refit = robust_affine(xr, yr, ud, vd, acx, acy, kAnchorRefitMinPoints, 1.0, aff);
            }
            if (refit) {
                have_affine = true;
            } else if (gm.locked && !seeds.model_rejected) {
                have_affine = true;
                std::copy(gm.a, gm.a + 6, aff);
                acx = gm.cx;
                acy = gm.cy;
            }
        };
        seed_lattice();
        // the plain lattice seeds the field when its anchors converge (the
        // output gate at 90% of those attempted) and it reaches SPARSE
        // quality or better; else the CNN lattice takes over
        int converged = 0;
        for (const AnchorRec& a : seeds.an) converged += a.output ? 1 : 0;
        const bool plain_seeds = seeds.quality >= 1 && converged >= 0.9f * (float)std::max(1, seeds.attempted);
        if (cnn_ok && !use_cnn && !plain_seeds) {
            use_cnn = true;
            plan();
            seeds = Seeds{};
            seeds.gU = (float)gm.t_u;
            seeds.gV = (float)gm.t_v;
            seed_lattice();
        }
        // publish_anchor_results (full_field_anchors.cpp:576-607)
        int anchor_pts = 0;
        {
            std::vector<int> nodes;
            std::vector<SolveOut> res;
            for (const AnchorRec& a : seeds.an) {
                if (!a.output || !a.kept) continue;
                SolveOut o = {};
                std::copy(a.p, a.p + 6, o.p);
                o.corr = a.corr;
                o.status = 0;
                o.iters = a.iters;
                o.simplex = a.simplex;
                o.init = 1;
                nodes.push_back(a.node);
                res.push_back(o);
                solved_h[a.node] = 1;
                anchor_pts++;
            }
            if (!nodes.empty()) {
                if (!nodes_d_.reserve(nodes.size()) || !out_d_.reserve(res.size())) return kErrInit;
                ck(hipMemcpyAsync(nodes_d_.p, nodes.data(), sizeof(int) * nodes.size(), hipMemcpyHostToDevice, st_));
                ck(hipMemcpyAsync(out_d_.p, res.data(), sizeof(SolveOut) * res.size(), hipMemcpyHostToDevice, st_));
                ck(launch_apply(g, nodes_d_.p, out_d_.p, (int)nodes.size(), kMeshAnchor, false, st_));
            }
        }
        ticker.set(anchor_pts);
        tm_.seeding = ms_since(t_seed);
        if (cancel.requested()) return kErrCancelled;

        float gU = seeds.gU, gV = seeds.gV;
        const bool have_seed = seeds.solved > 0 || seeds.accepted >= 3;
        const bool exec_c = seeds.quality == 0 && !have_seed;

        // ---- guess field (build_mesh_guess_field) --------------------------
        const auto t_mesh = clk::now();
        std::vector<TriDev> tris;
        std::vector<int> cell_tri;
        if (LW >= 2 && LH >= 2) {
            std::vector<unsigned char> kept(seeds.an.size());
            std::vector<float> rx(kept.size()), ry(kept.size()), uu(kept.size()), vv(kept.size());
            for (size_t i = 0; i < kept.size(); ++i) {
                const AnchorRec& a = seeds.an[i];
                kept[i] = a.kept && a.rx > 0 && a.rx < w_ && a.ry > 0 && a.ry < h_;
                rx[i] = a.rx;
                ry[i] = a.ry;
                uu[i] = a.p[0];
                vv[i] = a.p[1];
            }
            lattice_triangles(LW, LH, kept, rx, ry, uu, vv, tris, cell_tri);
        } else {
            cell_tri.assign(1, 0);
        }
        std::vector<int> lat(lx);
        lat.insert(lat.end(), ly.begin(), ly.end());
        if (!tris_d_.reserve(std::max<size_t>(1, tris.size())) || !celltri_d_.reserve(cell_tri.size()) ||
            !lat_d_.reserve(lat.size()))
            return kErrInit;
        if (!tris.empty())
            ck(hipMemcpyAsync(tris_d_.p, tris.data(), sizeof(TriDev) * tris.size(), hipMemcpyHostToDevice, st_));
        ck(hipMemcpyAsync(celltri_d_.p, cell_tri.data(), sizeof(int) * cell_tri.size(), hipMemcpyHostToDevice, st_));
        ck(hipMemcpyAsync(lat_d_.p, lat.data(), sizeof(int) * lat.size(), hipMemcpyHostToDevice, st_));
        GuessArgs ga = {};
        ga.tris = tris_d_.p;
        ga.cell_tri = celltri_d_.p;
        ga.lat_x = lat_d_.p;
        ga.lat_y = lat_d_.p + LW;
        ga.LW = LW;
        ga.LH = LH;
        ga.quality = seeds.quality;
        ga.have_affine = have_affine ? 1 : 0;
        std::copy(aff, aff + 6, ga.aff);
        ga.acx = acx;
        ga.acy = acy;
        ga.gU = seeds.gU;
        ga.gV = seeds.gV;
        ga.gw = gw;
        ga.gh = gh;
        ga.rect_x = p.rect_x;
        ga.rect_y = p.rect_y;
        ga.step = p.step;
        float* gu = guess_.p;
        ck(launch_guess(ga, g, in_mesh, gu, gu + NN, gu + 2 * NN, gu + 3 * NN, gu + 4 * NN, gu + 5 * NN, tmp_.p, st_));
        ck(hipStreamSynchronize(st_));
        tm_.mesh = ms_since(t_mesh);

        // ---- Path C (full_field_path_c.cpp:27-156) --------------------------
        int pcx = gw / 2, pcy = gh / 2;
        if (exec_c) {
            std::vector<int> cand;
            if (!seeds.ref_x.empty()) {
                const float gcx = p.rect_x + (gw / 2.f) * p.step, gcy = p.rect_y + (gh / 2.f) * p.step;
                float md = 1e9f;
                int bi = -1;
                for (size_t i = 0; i < seeds.ref_x.size(); ++i) {
                    const float dx = seeds.ref_x[i] - gcx, dy = seeds.ref_y[i] - gcy;
                    if (dx * dx + dy * dy < md) {
                        md = dx * dx + dy * dy;
                        bi = (int)i;
                    }
                }
                if (bi >= 0) {
                    const int cx = std::max(0, std::min(gw - 1, (int)std::round((seeds.ref_x[bi] - p.rect_x) / p.step)));
                    const int cy = std::max(0, std::min(gh - 1, (int)std::round((seeds.ref_y[bi] - p.rect_y) / p.step)));
                    if (!solved_h[(size_t)cy * gw + cx]) cand.push_back(cy * gw + cx);
                }
            }
            {
                const float ccx = gw / 2.0f, ccy = gh / 2.0f;
                std::vector<std::pair<float, int>> fr;
                for (int y = 0; y < gh; ++y)
                    for (int x = 0; x < gw; ++x) {
                        if (solved_h[(size_t)y * gw + x]) continue;
                        const float dx = x - ccx, dy = y - ccy;
                        fr.emplace_back(dx * dx + dy * dy, y * gw + x);
                    }
                std::stable_sort(fr.begin(), fr.end(), [](const std::pair<float, int>& a, const std::pair<float, int>& b) {
                    return a.first < b.first;
                });
                for (const auto& nd : fr) {
                    if ((int)cand.size() >= kPathCMaxCandidates) break;
                    if (!cand.empty() && cand.front() == nd.second) continue;
                    cand.push_back(nd.second);
                }
            }
            if (cand.empty()) return kErrRoi;
            int rc = kErrSeed;
            for (int flat : cand) {
                if (cancel.requested()) return kErrCancelled;
                Task t = {};
                t.cx = p.rect_x + (flat % gw) * p.step;
                t.cy = p.rect_y + (flat / gw) * p.step;
                t.pool = flat;
                t.mode = AUTO_SEARCH;
                t.g[0] = gU;
                t.g[1] = gV;
                t.node = flat;
                std::vector<SolveOut> r1;
                const auto t0 = clk::now();
                if (solve_tasks({t}, r1, p.use6, p.subset, tol) != 0) return kErrInit;
                icgn_ms_est += ms_since(t0);
                if (!r1[0].init) continue;
                if (r1[0].status == 0 && r1[0].corr <= kCorrAccept) {
                    pcx = flat % gw;
                    pcy = flat / gw;
                    gU = r1[0].p[0];
                    gV = r1[0].p[1];
                    rc = 0;
                    break;
                }
            }
            if (rc != 0) {
                if (!gm.locked) return rc;
                pcx = cand.front() % gw;
                pcy = cand.front() / gw;
            }
        }

        // ---- Path A: the grid launch ----------------------------------------
        const auto t_a = clk::now();
        {
            if (!stat_slots_.reserve((size_t)kStatSlots * kStatStride)) throw HipFail{hipErrorOutOfMemory};
            ck(hipMemsetAsync(counters_.p, 0, sizeof(int32_t) * 8, st_));
            ck(launch_zero(stat_slots_.p, kStatSlots * kStatStride, st_));
            SolveArgs a = {};
            a.nitems = N;
            a.bail = bail_.p;
            a.nbail = counters_.p + 8;
            a.stats = stat_slots_.p;
            a.grid.grid = g;
            a.grid.in_mesh = in_mesh;
            a.grid.gu = gu;
            a.grid.gv = gu + NN;
            a.grid.gux = gu + 2 * NN;
            a.grid.guy = gu + 3 * NN;
            a.grid.gvx = gu + 4 * NN;
            a.grid.gvy = gu + 5 * NN;
            a.grid.gw = gw;
            a.grid.gh = gh;
            a.grid.rect_x = p.rect_x;
            a.grid.rect_y = p.rect_y;
            a.grid.step = p.step;
            a.grid.w = w_;
            a.grid.h = h_;
            // learned seeding: a node the lattice does not cover has only the
            // global model's guess, which the field's bumps can leave far
            // behind; the layered flood claims it from its solved neighbours
            a.grid.skip_model = use_cnn ? 1 : 0;
            // and a node IC-GN cannot finish from the lattice's guess is claimed
            // again by the flood from a solved neighbour, so no rescue here
            a.grid.no_rescue = use_cnn ? 1 : 0;
            a.pool = pool_.p;
            a.refI = I_.p;
            a.refGx = gx_.p;
            a.refGy = gy_.p;
            set_precision_args(a);
            set_def_args(a);
            a.w = w_;
            a.h = h_;
            a.lm = 1;
            a.tol = tol;
            // one launch per subset size (one in all unless sizes are mixed),
            // each in the shape the tuner gives it and timed on its own; with
            // several sizes, the classes after the first run at once on streams
            // of their own, forked from the engine's stream and joined back
            // into it, each with its own hand-off list (its nodes' share of
            // the bail array) and counter, so that one class's slow subsets
            // overlap the others' launches
            const size_t nc = classes_.size();
            ensure_class_streams(nc);
            std::vector<std::vector<int>> keys(nc);
            std::vector<int> shapes(nc);
            std::vector<SolveArgs> cargs(nc);
            hipEvent_t ev_fork = class_ev_[2 * nc], ev_done = class_ev_[2 * nc + 1];
            ck(hipEventRecord(ev_fork, st_));
            // every class's fast path first, then every class's full solver:
            // streams that share a hardware queue then never hold a fast path
            // behind another class's slow subsets
            for (size_t ci = 0; ci < nc; ++ci) {
                const DimClass& c = classes_[ci];
                const bool listed = c.off >= 0;
                hipStream_t cs = ci == 0 ? st_ : class_st_[ci - 1];
                if (ci > 0) ck(hipStreamWaitEvent(cs, ev_fork, 0));
                keys[ci] = {gw, gh, c.dim, p.use6 ? 1 : 0, listed ? 1 : 0, c.n,
                            (def_f32_ready_ ? 1 : 0) + (def_coef_ready_ ? 2 : 0)};
                shapes[ci] = shape_for(keys[ci], listed);
                SolveArgs ac = a;
                ac.items = listed ? class_nodes_d_.p + c.off : nullptr;
                ac.nitems = c.n;
                ac.dim = c.dim;
                ac.grid.dim = c.dim;
                ac.bail = bail_.p + (listed ? c.off : 0);
                ac.nbail = ci == 0 ? counters_.p + 8 : class_nbail_.p + ci;
                ac.wpb = kShapes[shapes[ci]].wpb;
                ac.tile = listed ? 0 : kShapes[shapes[ci]].tile;
                cargs[ci] = ac;
                ck(hipEventRecord(class_ev_[2 * ci], cs));
                ck(launch_solve(ac, p.use6, def_f32_ready_, cs, kSolveFast));
            }
            for (size_t ci = 0; ci < nc; ++ci) {
                hipStream_t cs = ci == 0 ? st_ : class_st_[ci - 1];
                ck(launch_solve(cargs[ci], p.use6, def_f32_ready_, cs, kSolveFull));
                ck(hipEventRecord(class_ev_[2 * ci + 1], cs));
            }
            for (size_t ci = 1; ci < nc; ++ci) ck(hipStreamWaitEvent(st_, class_ev_[2 * ci + 1], 0));
            ck(hipEventRecord(ev_done, st_));
            int sa[9];
            std::vector<int32_t> handed(nc, 0);
            ck(launch_sum_stats(stat_slots_.p, counters_.p, st_));
            ck(hipMemcpyAsync(sa, counters_.p, sizeof(sa), hipMemcpyDeviceToHost, st_));
            if (nc > 1)
                ck(hipMemcpyAsync(handed.data() + 1, class_nbail_.p + 1, sizeof(int32_t) * (nc - 1),
                                  hipMemcpyDeviceToHost, st_));
            ck(hipStreamSynchronize(st_));
            float kms = 0;
            ck(hipEventElapsedTime(&kms, ev_fork, ev_done));
            for (size_t ci = 0; ci < nc; ++ci) {
                float cms = 0;
                ck(hipEventElapsedTime(&cms, class_ev_[2 * ci], class_ev_[2 * ci + 1]));
                shape_timed(keys[ci], classes_[ci].off >= 0, shapes[ci], cms);
            }
            for (size_t ci = 1; ci < nc; ++ci) sa[8] += handed[ci];
            tm_.kernel_path_a = kms;
            tm_.path_a_items = N;
            launch_log.emplace_back((int)N, sa[8]);
            if (node_report) {
                after_a_.resize(NN);
                ck(hipMemcpy(after_a_.data(), g.solved, NN, hipMemcpyDeviceToHost));
                report_planes_a.resize(7 * NN);
                ck(hipMemcpy(report_planes_a.data(), gplanes_.p, sizeof(float) * 7 * NN, hipMemcpyDeviceToHost));
                report_guess.resize(6 * NN);
                ck(hipMemcpy(report_guess.data(), guess_.p, sizeof(float) * 6 * NN, hipMemcpyDeviceToHost));
            }
            st_a.points = sa[0];
            st_a.iters = sa[1];
            st_a.calls = sa[2];
            st_a.saved = sa[3];
            st_a.dead = sa[4];
            tm_.path_a_iters = sa[1];
        }
        tm_.path_a = ms_since(t_a);
        ticker.set(anchor_pts + st_a.points);
        if (cancel.requested()) return kErrCancelled;

        // ---- Path B (full_field_path_b.cpp:148-412) ---------------------------
        const auto t_b = clk::now();
        int open_nodes = 0;
        ck(launch_count_open(g, N, counters_.p + 13, st_));
        ck(hipMemcpyAsync(&open_nodes, counters_.p + 13, sizeof(int), hipMemcpyDeviceToHost, st_));
        ck(hipStreamSynchronize(st_));
        // with every node solved or skipped there is nothing to claim, and no global seed to try
        if (open_nodes > 0) {
            // the accepted flags (solved with a real corr), made on the device into
            // the valid-flag buffer, which the end of the run rewrites; the host's
            // floods read them and the node planes through page-locked staging
            if (!pin_planes_.reserve(7 * NN) || !pin_flags_.reserve(2 * NN)) throw HipFail{hipErrorOutOfMemory};
            float* const hp = pin_planes_.p;
            const uint8_t* const hsolved = pin_flags_.p;
            const uint8_t* const haccepted = pin_flags_.p + NN;
            ck(launch_valid(g, valid_d, N, st_));
            bool fetched = false;
            const auto fetch = [&]() {
                if (fetched) return;
                ck(hipMemcpyAsync(hp, gplanes_.p, sizeof(float) * 7 * NN, hipMemcpyDeviceToHost, st_));
                ck(hipMemcpyAsync(pin_flags_.p, g.solved, NN, hipMemcpyDeviceToHost, st_));
                ck(hipMemcpyAsync(pin_flags_.p + NN, valid_d, NN, hipMemcpyDeviceToHost, st_));
                ck(hipStreamSynchronize(st_));
                fetched = true;
            };
            const float* hu = hp;
            const float* hv = hu + NN;
            const float* hux = hv + NN;
            const float* huy = hux + NN;
            const float* hvx = huy + NN;
            const float* hvy = hvx + NN;
            const float* hcorr = hvy + NN;
            const int dx4[4] = {1, -1, 0, 0}, dy4[4] = {0, 0, 1, -1};
            bool flooded = false;
            const int reach = layered_path_b > 0 ? layered_path_b : use_cnn ? kLayeredReachLearned : 0;
            if (reach > 0 && flood_on_device && reach <= kFloodDeviceMaxReach && grid_dims_.empty()) {
                // Layered Path B with its frontier on the device (flood.hpp): the
                // same layers, claims and seeds as the host's below, with two
                // small read-backs per layer (the source and task counts)
                const FloodDev f = flood_dev(gw, gh);
                ck(launch_flood_init(f, valid_d, g.solved, gplanes_.p, reach, st_));
                int32_t* const pc = pin_fl_.p;
                int pending = 0;   // the last layer's task count, its hand-off count still to read
                for (int layer = 0;; ++layer) {
                    if (cancel.requested()) return kErrCancelled;
                    ck(launch_flood_sources(f, st_));
                    ck(hipMemcpyAsync(pc, f.counters, sizeof(int32_t) * kFcCount, hipMemcpyDeviceToHost, st_));
                    ck(hipMemcpyAsync(pc + kFcCount, counters_.p + 8, sizeof(int32_t), hipMemcpyDeviceToHost, st_));
                    ck(hipStreamSynchronize(st_));
                    if (pending > 0) launch_log.emplace_back(pending, pc[kFcCount]);
                    pending = 0;
                    ticker.set(anchor_pts + st_a.points + pc[kFcPoints]);
                    const int ns = pc[kFcSources];
                    if (ns == 0) break;
                    flooded = true;
                    ck(launch_flood_sort(f, ns, fl_sort_temp_.p, fl_sort_temp_.n, st_));
                    ck(launch_flood_search(f, ns, reach, layer + 1, p.rect_x, p.rect_y, p.step, use_cnn, st_));
                    ck(hipMemcpyAsync(pc, f.counters, sizeof(int32_t) * kFcCount, hipMemcpyDeviceToHost, st_));
                    ck(hipStreamSynchronize(st_));
                    const int nt = pc[kFcTasks];
                    if (nt == 0) break;
                    if (solve_tasks_device(f.tasks, nt, fl_res_.p, p.use6, p.subset, tol) != 0) return kErrInit;
                    tm_.path_b_launches++;
                    tm_.path_b_miss += nt;
                    ck(launch_flood_update(f, fl_res_.p, nt, st_));
                    pending = nt;
                }
                // one entry per claimed node, holding its last claim
                ck(launch_flood_commit_list(f, st_));
                ck(hipMemcpyAsync(pc, f.counters, sizeof(int32_t) * kFcCount, hipMemcpyDeviceToHost, st_));
                ck(hipStreamSynchronize(st_));
                st_b.points += pc[kFcPoints];
                st_b.iters += pc[kFcIters];
                st_b.calls += pc[kFcCalls];
                st_b.saved += pc[kFcSaved];
                st_b.dead += pc[kFcDead];
                if (pc[kFcCommitted] > 0) {
                    ck(launch_apply(g, f.commit_nodes, f.commit_list, pc[kFcCommitted], -1, true, st_));
                    ck(hipStreamSynchronize(st_));
                }
                ticker.set(anchor_pts + st_a.points + st_b.points);
            } else if (reach > 0) {
                fetch();
                // Layered Path B. ok: accepted (solved before Path B with a real corr, or
                // accepted here), w: their warps (the seven planes of hp, updated in
                // place: the reference flood below runs only when no layer ran); fin: solved,
                // skipped or finally failed; tried: claims must come closer than this
                // (after a failed far claim only an adjacent one is made, so a node is
                // claimed at most twice). Under learned seeding no claim takes the
                // Simplex rescue: an adjacent claim's carried guess is off by the
                // field's curvature over one step, and the nodes that fail it are
                // mostly those whose subsets leave the frame.
                // (node-sized scratch kept on the engine: refilled, not reallocated)
                FloodScratch& fs = flood_;
                fs.ok.assign(haccepted, haccepted + NN);
                fs.fin.assign(hsolved, hsolved + NN);
                fs.tried.assign(NN, (int8_t)(reach + 1));
                // dist and src are written before they are read (for each node a
                // layer's search reaches); stamps are compared against this run's
                // epoch plus the layer, so neither needs clearing between runs
                if (fs.stamp.size() != NN || fs.epoch > (1 << 30)) {
                    fs.stamp.assign(NN, -1);
                    fs.epoch = 0;
                }
                fs.dist.resize(NN);
                fs.src.resize(NN);
                fs.commit_at.assign(NN, -1);
                const int epoch = fs.epoch;
                std::vector<uint8_t>& ok = fs.ok;
                std::vector<uint8_t>& fin = fs.fin;
                std::vector<int8_t>& tried = fs.tried;
                std::vector<int>& stamp = fs.stamp;
                std::vector<int>& dist = fs.dist;
                std::vector<int>& src = fs.src;
                std::vector<int>& commit_at = fs.commit_at;
                float* const w = hp;
                std::vector<int> open, bfs, sources;
                // the open nodes, eight flags at a time (most are solved)
                {
                    size_t i = 0;
                    for (; i + 8 <= NN; i += 8) {
                        uint64_t f8;
                        std::memcpy(&f8, hsolved + i, 8);
                        if (f8 == 0x0101010101010101ull) continue;
                        for (size_t k = i; k < i + 8; ++k)
                            if (!hsolved[k]) open.push_back((int)k);
                    }
                    for (; i < NN; ++i)
                        if (!hsolved[i]) open.push_back((int)i);
                }
                // one entry per node, holding its last claim: a node claimed twice
                // (a failed far claim, then an adjacent one) must not reach the
                // apply kernel twice, where the two writes would race
                std::vector<int> commit_nodes;
                std::vector<SolveOut> commit_res;
                for (int layer = 0; !open.empty(); ++layer) {
                    fs.epoch = epoch + layer + 1;
                    if (cancel.requested()) return kErrCancelled;
                    // the accepted nodes next to open ones, in the order the flood pops them
                    sources.clear();
                    for (int i : open) {
                        const int x = i % gw, y = i / gw;
                        for (int k = 0; k < 4; ++k) {
                            const int nx = x + dx4[k], ny = y + dy4[k];
                            if (nx < 0 || nx >= gw || ny < 0 || ny >= gh) continue;
                            const int j = ny * gw + nx;
                            if (ok[j] && stamp[j] != epoch + layer) {
                                stamp[j] = epoch + layer;
                                sources.push_back(j);
                            }
                        }
                    }
                    if (sources.empty()) break;
                    flooded = true;
                    std::sort(sources.begin(), sources.end(), [&](int a, int b) {
                        const float ca = w[6 * NN + a], cb = w[6 * NN + b];
                        return ca != cb ? ca < cb : a < b;
                    });
                    bfs = sources;
                    for (int j : sources) {
                        dist[j] = 0;
                        src[j] = j;
                    }
                    for (size_t head = 0; head < bfs.size(); ++head) {
                        const int j = bfs[head];
                        if (dist[j] >= reach) continue;
                        const int x = j % gw, y = j / gw;
                        for (int k = 0; k < 4; ++k) {
                            const int nx = x + dx4[k], ny = y + dy4[k];
                            if (nx < 0 || nx >= gw || ny < 0 || ny >= gh) continue;
                            const int m = ny * gw + nx;
                            if (fin[m] || ok[m] || stamp[m] == epoch + layer) continue;
                            stamp[m] = epoch + layer;
                            dist[m] = dist[j] + 1;
                            src[m] = src[j];
                            bfs.push_back(m);
                        }
                    }
                    std::vector<Task> tasks;
                    std::vector<int> tdist;
                    for (size_t k = sources.size(); k < bfs.size(); ++k) {
                        const int i = bfs[k];
                        if (dist[i] >= tried[i]) continue;
                        const int s0 = src[i], x = i % gw, y = i / gw;
                        const float ddx = (float)((x - s0 % gw) * p.step), ddy = (float)((y - s0 / gw) * p.step);
                        Task tk = {};
                        tk.cx = p.rect_x + x * p.step;
                        tk.cy = p.rect_y + y * p.step;
                        tk.pool = i;
                        tk.node = i;
                        tk.mode = dist[i] == 1 && !use_cnn ? NO_SEARCH : NO_SIMPLEX;
                        tk.g[0] = w[s0] + w[2 * NN + s0] * ddx + w[3 * NN + s0] * ddy;
                        tk.g[1] = w[NN + s0] + w[4 * NN + s0] * ddx + w[5 * NN + s0] * ddy;
                        for (int c = 0; c < 4; ++c) tk.g[2 + c] = w[(2 + c) * NN + s0];
                        tasks.push_back(tk);
                        tdist.push_back(dist[i]);
                    }
                    if (tasks.empty()) break;
                    std::vector<SolveOut> res;
                    if (solve_tasks(tasks, res, p.use6, p.subset, tol) != 0) return kErrInit;
                    tm_.path_b_launches++;
                    tm_.path_b_miss += (int)tasks.size();
                    for (size_t t = 0; t < tasks.size(); ++t) {
                        const SolveOut& r = res[t];
                        const int i = tasks[t].node;
                        if (r.init) {
                            st_b.record(r);
                            if (commit_at[i] < 0) {
                                commit_at[i] = (int)commit_nodes.size();
                                commit_nodes.push_back(i);
                                commit_res.push_back(r);
                            } else {
                                commit_res[commit_at[i]] = r;
                            }
                        }
                        if (r.init && r.status == 0 && r.corr <= kCorrAccept) {
                            st_b.points++;
                            ok[i] = 1;
                            fin[i] = 1;
                            for (int c = 0; c < 6; ++c) w[c * NN + i] = r.p[c];
                            w[6 * NN + i] = r.corr;
                        } else if (tdist[t] == 1 || !r.init) {
                            fin[i] = 1;
                        } else {
                            tried[i] = 2;   // a failed far claim waits for an adjacent one
                        }
                    }
                    size_t kept = 0;
                    for (int i : open)
                        if (!fin[i]) open[kept++] = i;
                    open.resize(kept);
                    ticker.set(anchor_pts + st_a.points + st_b.points);
                }
                if (!commit_nodes.empty()) {
                    const int n = (int)commit_nodes.size();
                    if (!nodes_d_.reserve(n) || !out_d_.reserve(n)) throw HipFail{hipErrorOutOfMemory};
                    ck(hipMemcpyAsync(nodes_d_.p, commit_nodes.data(), sizeof(int) * n, hipMemcpyHostToDevice, st_));
                    ck(hipMemcpyAsync(out_d_.p, commit_res.data(), sizeof(SolveOut) * n, hipMemcpyHostToDevice, st_));
                    ck(launch_apply(g, nodes_d_.p, out_d_.p, n, -1, true, st_));
                    ck(hipStreamSynchronize(st_));
                }
            }
            // the reference's flood; the layered one hands over only when no node was
            // accepted before it (the flood's global seeds start it)
            if (!flooded) {
                fetch();
                struct SeedNode {
                    int x, y;
                    float u, v, ux, uy, vx, vy, corr;
                };
                struct Rel {
                    bool operator()(const SeedNode& a, const SeedNode& b) const {
                        if (a.corr != b.corr) return a.corr > b.corr;
                        if (a.y != b.y) return a.y > b.y;
                        return a.x > b.x;
                    }
                };
                std::vector<unsigned char> claimed(hsolved, hsolved + NN);
                std::vector<SeedNode> boundary, global;
                const SeedNode pc{pcx, pcy, gU, gV, 0, 0, 0, 0, 0};
                bool pc_queued = false;
                for (int y = 0; y < gh; ++y)
                    for (int x = 0; x < gw; ++x) {
                        const size_t i = (size_t)y * gw + x;
                        if (!hsolved[i] || hcorr[i] < 0.f) continue;
                        bool touch = false;
                        for (int k = 0; k < 4; ++k) {
                            const int nx = x + dx4[k], ny = y + dy4[k];
                            if (nx >= 0 && nx < gw && ny >= 0 && ny < gh && !hsolved[(size_t)ny * gw + nx]) {
                                touch = true;
                                break;
                            }
                        }
                        if (touch) boundary.push_back({x, y, hu[i], hv[i], hux[i], huy[i], hvx[i], hvy[i], hcorr[i]});
                    }
                if (boundary.empty()) {
                    struct C {
                        int ix, iy;
                        float u, v, dist, mag;
                    };
                    std::vector<C> cands;
                    const float gcx = p.rect_x + (gw / 2.f) * p.step, gcy = p.rect_y + (gh / 2.f) * p.step;
                    for (size_t i = 0; i < seeds.ref_x.size(); ++i) {
                        const float fx = seeds.ref_x[i], fy = seeds.ref_y[i];
                        const int ix = (int)std::round((fx - p.rect_x) / (float)p.step);
                        const int iy = (int)std::round((fy - p.rect_y) / (float)p.step);
                        if (ix < 0 || ix >= gw || iy < 0 || iy >= gh || hsolved[(size_t)iy * gw + ix]) continue;
                        const float du = seeds.def_x[i] - fx, dv = seeds.def_y[i] - fy;
                        const float wx = p.rect_x + ix * p.step, wy = p.rect_y + iy * p.step;
                        const float dc = std::sqrt((wx - gcx) * (wx - gcx) + (wy - gcy) * (wy - gcy));
                        cands.push_back({ix, iy, du, dv, dc, std::sqrt(du * du + dv * dv)});
                    }
                    if (cands.empty()) {
                        global.push_back(pc);
                        pc_queued = true;
                    } else {
                        std::sort(cands.begin(), cands.end(), [](const C& a, const C& b) {
                            const float ma = std::floor(a.mag), mb = std::floor(b.mag);
                            if (ma != mb) return ma < mb;
                            if (a.dist != b.dist) return a.dist < b.dist;
                            if (a.iy != b.iy) return a.iy < b.iy;
                            return a.ix < b.ix;
                        });
                        for (size_t k = 0; k < cands.size() && k < 5; ++k)
                            global.push_back({cands[k].ix, cands[k].iy, cands[k].u, cands[k].v, 0, 0, 0, 0, 0});
                    }
                }
                std::priority_queue<SeedNode, std::vector<SeedNode>, Rel> q;
                for (const auto& b : boundary) q.push(b);
                struct Cell {
                    Task t;
                    SolveOut r;
                    int from;   // flat index of the node whose warp seeded the guess (-1: a global seed)
                };
                // commit one solved cell: host mirror, queue, counters
                std::vector<int> commit_nodes;
                std::vector<SolveOut> commit_res;
                auto commit = [&](const Cell& c) {
                    if (!c.r.init) return;
                    st_b.record(c.r);
                    commit_nodes.push_back(c.t.node);
                    commit_res.push_back(c.r);
                    if (c.r.status == 0 && c.r.corr <= kCorrAccept) {
                        st_b.points++;
                        q.push({c.t.node % gw, c.t.node / gw, c.r.p[0], c.r.p[1], c.r.p[2], c.r.p[3], c.r.p[4], c.r.p[5],
                                c.r.corr});
                    }
                };
                auto make_task = [&](int x, int y, const float* gs) {
                    Task t = {};
                    t.cx = p.rect_x + x * p.step;
                    t.cy = p.rect_y + y * p.step;
                    t.pool = y * gw + x;
                    t.mode = NO_SEARCH;
                    std::copy(gs, gs + 6, t.g);
                    t.node = t.pool;
                    return t;
                };
                // the guess a popped node hands to its neighbour: its warp carried by the grid offset
                auto seed_task = [&](const SeedNode& cur, int nx, int ny) {
                    const float ddx = (float)((nx - cur.x) * p.step), ddy = (float)((ny - cur.y) * p.step);
                    const float gs[6] = {cur.u + cur.ux * ddx + cur.uy * ddy, cur.v + cur.vx * ddx + cur.vy * ddy,
                                         cur.ux, cur.uy, cur.vx, cur.vy};
                    return make_task(nx, ny, gs);
                };
                // Speculative first claims. The queue pops the nodes solved before Path B in
                // (corr, y, x) order, and pushes never reorder them, so an unsolved node is
                // claimed by the first of its pre-solved neighbours in that order unless a node
                // solved inside Path B claims it earlier. All such first claims are solved in
                // one launch here; the round replay below takes their results when the claim
                // matches and solves only the other claims. Each task's result depends on the
                // task alone, so the replay commits exactly what round-by-round solving would.
                std::vector<int> spec_from(NN, -1);
                std::vector<SolveOut> spec_res;
                std::vector<int> spec_slot(NN, -1);
                {
                    std::vector<Task> spec;
                    for (int y = 0; y < gh && speculate_path_b; ++y)
                        for (int x = 0; x < gw; ++x) {
                            const int i = y * gw + x;
                            if (claimed[i]) continue;
                            int best = -1;
                            for (int k = 0; k < 4; ++k) {
                                const int nx = x + dx4[k], ny = y + dy4[k];
                                if (nx < 0 || nx >= gw || ny < 0 || ny >= gh) continue;
                                const int j = ny * gw + nx;
                                if (!hsolved[j] || hcorr[j] < 0.f) continue;
                                if (best < 0) {
                                    best = j;
                                    continue;
                                }
                                const int bx = best % gw, by = best / gw;
                                const bool better = hcorr[j] != hcorr[best] ? hcorr[j] < hcorr[best]
                                                    : ny != by           ? ny < by
                                                                         : nx < bx;
                                if (better) best = j;
                            }
                            if (best < 0) continue;
                            const SeedNode m{best % gw, best / gw, hu[best], hv[best], hux[best], huy[best],
                                             hvx[best], hvy[best], hcorr[best]};
                            spec_from[i] = best;
                            spec_slot[i] = (int)spec.size();
                            spec.push_back(seed_task(m, x, y));
                        }
                    if (!spec.empty() && solve_tasks(spec, spec_res, p.use6, p.subset, tol) != 0) return kErrInit;
                    tm_.path_b_spec = (int)spec.size();
                }
                auto flush_commits = [&]() {
                    if (commit_nodes.empty()) return;
                    const int n = (int)commit_nodes.size();
                    if (!nodes_d_.reserve(n) || !out_d_.reserve(n)) throw HipFail{hipErrorOutOfMemory};
                    ck(hipMemcpyAsync(nodes_d_.p, commit_nodes.data(), sizeof(int) * n, hipMemcpyHostToDevice, st_));
                    ck(hipMemcpyAsync(out_d_.p, commit_res.data(), sizeof(SolveOut) * n, hipMemcpyHostToDevice, st_));
                    ck(launch_apply(g, nodes_d_.p, out_d_.p, n, -1, true, st_));
                    ck(hipStreamSynchronize(st_));
                    commit_nodes.clear();
                    commit_res.clear();
                };
                for (const auto& s : global) {
                    const int flat = s.y * gw + s.x;
                    if (claimed[flat]) continue;
                    claimed[flat] = 1;
                    const float gs[6] = {s.u, s.v, 0, 0, 0, 0};
                    Cell c{make_task(s.x, s.y, gs), {}, -1};
                    std::vector<SolveOut> r1;
                    if (solve_tasks({c.t}, r1, p.use6, p.subset, tol) != 0) return kErrInit;
                    c.r = r1[0];
                    commit(c);
                }
                if (q.empty() && !global.empty() && !pc_queued) {
                    const int flat = pc.y * gw + pc.x;
                    if (pc.x >= 0 && pc.x < gw && pc.y >= 0 && pc.y < gh && !claimed[flat]) {
                        claimed[flat] = 1;
                        const float gs[6] = {pc.u, pc.v, 0, 0, 0, 0};
                        Cell c{make_task(pc.x, pc.y, gs), {}, -1};
                        std::vector<SolveOut> r1;
                        if (solve_tasks({c.t}, r1, p.use6, p.subset, tol) != 0) return kErrInit;
                        c.r = r1[0];
                        commit(c);
                    }
                }
                struct Round {
                    std::vector<Cell> cells;
                    bool solved = false;
                };
                std::deque<Round> rounds;
                auto claim = [&]() -> bool {
                    if (q.empty()) return false;
                    Round r;
                    int popped = 0;
                    while (!q.empty() && popped < kPathBBatchNodes) {
                        const SeedNode cur = q.top();
                        q.pop();
                        ++popped;
                        for (int k = 0; k < 4; ++k) {
                            const int nx = cur.x + dx4[k], ny = cur.y + dy4[k];
                            if (nx < 0 || nx >= gw || ny < 0 || ny >= gh) continue;
                            const int flat = ny * gw + nx;
                            if (claimed[flat]) continue;
                            claimed[flat] = 1;
                            r.cells.push_back(Cell{seed_task(cur, nx, ny), {}, cur.y * gw + cur.x});
                        }
                    }
                    rounds.push_back(std::move(r));
                    return true;
                };
                while (rounds.size() < 2 && claim()) {
                }
                while (!rounds.empty()) {
                    if (cancel.requested()) return kErrCancelled;
                    if (!rounds.front().solved) {
                        // claims made by the speculated neighbour take its result; the rest are solved now
                        std::vector<Task> tasks;
                        std::vector<Cell*> miss;
                        for (Round& r : rounds)
                            if (!r.solved)
                                for (Cell& c : r.cells) {
                                    const int nd = c.t.node;
                                    if (c.from >= 0 && spec_from[nd] == c.from) {
                                        c.r = spec_res[spec_slot[nd]];
                                    } else {
                                        tasks.push_back(c.t);
                                        miss.push_back(&c);
                                    }
                                }
                        if (!tasks.empty()) {
                            std::vector<SolveOut> res;
                            if (solve_tasks(tasks, res, p.use6, p.subset, tol) != 0) return kErrInit;
                            for (size_t k = 0; k < miss.size(); ++k) miss[k]->r = res[k];
                            tm_.path_b_launches++;
                            tm_.path_b_miss += (int)tasks.size();
                        }
                        for (Round& r : rounds) r.solved = true;
                    }
                    for (const Cell& c : rounds.front().cells) commit(c);
                    rounds.pop_front();
                    while (rounds.size() < 2 && claim()) {
                    }
                    ticker.set(anchor_pts + st_a.points + st_b.points);
                }
                flush_commits();
            }
        }
        tm_.path_b = ms_since(t_b);
        if (cancel.requested()) return kErrCancelled;

        // ---- strain and packing (pack_full_field_output) ----------------------
        const auto t_s = clk::now();
        // median test: five passes, each from the previous pass's flags; a
        // pass after one that rejected nothing changes nothing, and its kernel
        // only copies the flags. Then VSG on the surviving nodes and the
        // row-major pack on the device, then one copy of the packed points.
        int* pass_fail = counters_.p + 16;   // one rejection count per pass
        int* pack_counts = counters_.p + 24; // written before the cut, dropped before the cut
        const int cap_points = out_capacity / 8;
        if (!pack_out_.reserve(8 * NN) || !pack_blocks_.reserve((size_t)pack_blocks(N)) || !pin_counts_.reserve(32))
            throw HipFail{hipErrorOutOfMemory};
        uint8_t* med = valid_d;
        {
            uint8_t* dst = ok_a;
            for (int pass = 0; pass < kOutlierMaxPasses; ++pass) {
                med = dst;
                dst = (dst == ok_a) ? ok_b : ok_a;
            }
        }
        float* exx = strain_.p;
        float* eyy = exx + NN;
        float* exy = eyy + NN;
        const std::vector<uint64_t> tail_key = {
            (uint64_t)gw, (uint64_t)gh, (uint64_t)p.rect_x, (uint64_t)p.rect_y, (uint64_t)p.step,
            (uint64_t)p.strain_window, (uint64_t)cap_points, (uint64_t)(uintptr_t)gplanes_.p,
            (uint64_t)(uintptr_t)gflags_.p, (uint64_t)(uintptr_t)strain_.p, (uint64_t)(uintptr_t)counters_.p,
            (uint64_t)(uintptr_t)pack_out_.p, (uint64_t)(uintptr_t)pack_blocks_.p, (uint64_t)(uintptr_t)pin_counts_.p,
            (uint64_t)(uintptr_t)giters_.p};
        run_chain(tail_chain_, tail_key, [&] {
            ck(launch_valid(g, valid_d, N, st_));
            ck(launch_zero(pass_fail, 16, st_));
            uint8_t* src = valid_d;
            uint8_t* dst = ok_a;
            for (int pass = 0; pass < kOutlierMaxPasses; ++pass) {
                ck(launch_outlier_pass(gw, gh, g.u, g.v, src, dst, pass > 0 ? pass_fail + pass - 1 : nullptr,
                                       pass_fail + pass, st_));
                src = dst;
                dst = (dst == ok_a) ? ok_b : ok_a;
            }
            ck(launch_vsg(gw, gh, p.step, p.strain_window, vsg_count(p.step, p.strain_window), g.u, g.v, src, exx,
                          eyy, exy, st_));
            ck(launch_pack(gw, gh, p.rect_x, p.rect_y, p.step, cap_points, src, g.u, g.v, exx, eyy, exy, g.corr,
                           pack_blocks_.p, pack_counts, pack_out_.p, st_));
            ck(hipMemcpyAsync(pin_counts_.p, counters_.p, sizeof(int32_t) * 32, hipMemcpyDeviceToHost, st_));
        });
        ck(hipStreamSynchronize(st_));
        int hc[32];
        std::memcpy(hc, pin_counts_.p, sizeof(hc));
        if (keep_field) {
            // the chain leaves the planes and the valid flags as they were
            last_dev_field = ComposeField{gplanes_.p, med, gw, gh, p.rect_x, p.rect_y, p.step};
            last_field = Field{};
            last_field.gw = gw;
            last_field.gh = gh;
            last_field.rect_x = p.rect_x;
            last_field.rect_y = p.rect_y;
            last_field.step = p.step;
            last_field.attempted = total_valid;
            std::vector<float> all(7 * NN);
            last_field.ok.resize(NN);
            last_field.med.resize(NN);
            ck(hipMemcpyAsync(all.data(), gplanes_.p, sizeof(float) * 7 * NN, hipMemcpyDeviceToHost, st_));
            ck(hipMemcpyAsync(last_field.ok.data(), valid_d, NN, hipMemcpyDeviceToHost, st_));
            ck(hipMemcpyAsync(last_field.med.data(), med, NN, hipMemcpyDeviceToHost, st_));
            ck(hipStreamSynchronize(st_));
            last_field.u.assign(all.begin(), all.begin() + NN);
            last_field.v.assign(all.begin() + NN, all.begin() + 2 * NN);
            last_field.ux.assign(all.begin() + 2 * NN, all.begin() + 3 * NN);
            last_field.uy.assign(all.begin() + 3 * NN, all.begin() + 4 * NN);
            last_field.vx.assign(all.begin() + 4 * NN, all.begin() + 5 * NN);
            last_field.vy.assign(all.begin() + 5 * NN, all.begin() + 6 * NN);
            last_field.corr.assign(all.begin() + 6 * NN, all.begin() + 7 * NN);
        }
        if (node_report) {
            std::vector<uint8_t> sol(NN), tg(NN);
            std::vector<float> corr(NN);
            ck(hipMemcpyAsync(sol.data(), g.solved, NN, hipMemcpyDeviceToHost, st_));
            ck(hipMemcpyAsync(tg.data(), g.tag, NN, hipMemcpyDeviceToHost, st_));
            ck(hipMemcpyAsync(corr.data(), g.corr, sizeof(float) * NN, hipMemcpyDeviceToHost, st_));
            report_planes.resize(7 * NN);
            ck(hipMemcpyAsync(report_planes.data(), gplanes_.p, sizeof(float) * 7 * NN, hipMemcpyDeviceToHost, st_));
            ck(hipStreamSynchronize(st_));
            report_tag = tg;
            report_path.assign(NN, kPathUnsolved);
            for (size_t i = 0; i < NN; ++i) {
                if (sol[i] && corr[i] < 0.0f) report_path[i] = kPathSkipped;
                else if (!sol[i]) report_path[i] = kPathUnsolved;
                else if (tg[i] == kMeshAnchor) report_path[i] = kPathAnchor;
                else if (i < after_a_.size() && after_a_[i]) report_path[i] = kPathA;
                else report_path[i] = kPathLater;
            }
        }
        int dropped_outlier = 0;
        for (int pass = 0; pass < kOutlierMaxPasses; ++pass) dropped_outlier += hc[16 + pass];
        const int valid = std::min(hc[24], cap_points);
        const int dropped_post = hc[25];
        if (valid > 0) {
            ck(hipMemcpyAsync(out, pack_out_.p, sizeof(float) * 8 * (size_t)valid,
                              out_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st_));
            ck(hipStreamSynchronize(st_));
        }
        tm_.strain = ms_since(t_s);
        ticker.stop();
        if (progress) {
            try {
                progress(100);
            } catch (...) {
            }
        }
        tm_.total = ms_since(t_total);
        tm_.icgn = icgn_ms_est + tm_.path_a;
        // fill_engine_metrics (full_field_solver_stats.cpp:182-280)
        const int ncopy = metrics_len < 16 ? 0 : metrics_len >= 24 ? 24 : metrics_len == 18 ? 17 : metrics_len;
        if (metrics && ncopy > 0) {
            const int iters = st_a.iters + st_b.iters + st_anchor.iters;
            float m[24];
            m[0] = (float)total_valid;
            m[1] = (float)valid;
            m[2] = (float)(total_valid - valid);
            m[3] = (float)st_a.points;
            m[4] = (float)st_b.points;
            m[5] = (float)(st_a.calls + st_b.calls + st_anchor.calls);
            m[6] = (float)(st_a.saved + st_b.saved + st_anchor.saved);
            m[7] = (float)(st_a.dead + st_b.dead + st_anchor.dead);
            m[8] = valid > 0 ? (float)iters / (float)valid : 0.0f;
            m[9] = (float)tm_.total;
            m[10] = (float)tm_.seeding;
            m[11] = (float)tm_.prepass;
            m[12] = (float)tm_.mesh;
            m[13] = (float)tm_.strain;
            m[14] = tm_.total > 0 ? (float)(valid / tm_.total) : 0.0f;
            m[15] = total_valid > 0 ? ((float)valid / total_valid) * 100.0f : 0.0f;
            m[16] = (float)seeds.quality;
            m[17] = (float)tm_.simplex;
            m[18] = (float)tm_.icgn;
            m[19] = (float)anchor_pts;
            m[20] = (float)dropped_post;
            m[21] = gm.locked ? 1.0f : 0.0f;
            m[22] = seeds.coverage;
            m[23] = (float)dropped_outlier;
            for (int i = 0; i < ncopy; ++i) metrics[i] = m[i];
        }
        return valid;
    } catch (const HipFail&) {
        return kErrInit;
    }
}

int Engine::strain_pack(const Field& f, int window, float* out, int out_capacity, int* dropped_post,
                        int* dropped_outlier) {
    const int gw = f.gw, gh = f.gh, N = gw * gh;
    const size_t NN = (size_t)N;
    if (N <= 0) return 0;
    try {
        if (!gplanes_.reserve(7 * NN) || !gflags_.reserve(7 * NN) || !strain_.reserve(3 * NN) || !counters_.reserve(32))
            return kErrInit;
        if (!pin_in_.reserve(2 * NN) || !pin_out_.reserve(3 * NN) || !pin_ok_in_.reserve(NN) ||
            !pin_ok_out_.reserve(NN) || !pin_counts_.reserve(32))
            return kErrInit;
        float* u = gplanes_.p;
        float* v = u + NN;
        uint8_t* ok0 = gflags_.p;
        uint8_t* ok1 = ok0 + NN;
        uint8_t* ok2 = ok1 + NN;
        float* exx = strain_.p;
        std::memcpy(pin_in_.p, f.u.data(), sizeof(float) * NN);
        std::memcpy(pin_in_.p + NN, f.v.data(), sizeof(float) * NN);
        std::memcpy(pin_ok_in_.p, f.ok.data(), NN);
        const std::vector<uint64_t> key = {
            (uint64_t)gw, (uint64_t)gh, (uint64_t)f.step, (uint64_t)window, (uint64_t)(uintptr_t)gplanes_.p,
            (uint64_t)(uintptr_t)gflags_.p, (uint64_t)(uintptr_t)strain_.p, (uint64_t)(uintptr_t)counters_.p,
            (uint64_t)(uintptr_t)pin_in_.p, (uint64_t)(uintptr_t)pin_out_.p, (uint64_t)(uintptr_t)pin_ok_in_.p,
            (uint64_t)(uintptr_t)pin_ok_out_.p, (uint64_t)(uintptr_t)pin_counts_.p};
        run_chain(strain_chain_, key, [&] {
            ck(hipMemcpyAsync(u, pin_in_.p, sizeof(float) * 2 * NN, hipMemcpyHostToDevice, st_));
            ck(hipMemcpyAsync(ok0, pin_ok_in_.p, NN, hipMemcpyHostToDevice, st_));
            // one rejection count per pass: a pass after one that rejected
            // nothing copies the flags
            int* pass_fail = counters_.p + 16;
            ck(launch_zero(pass_fail, kOutlierMaxPasses, st_));
            uint8_t* src = ok0;
            uint8_t* dst = ok1;
            for (int pass = 0; pass < kOutlierMaxPasses; ++pass) {
                ck(launch_outlier_pass(gw, gh, u, v, src, dst, pass > 0 ? pass_fail + pass - 1 : nullptr,
                                       pass_fail + pass, st_));
                src = dst;
                dst = (dst == ok1) ? ok2 : ok1;
            }
            ck(launch_vsg(gw, gh, f.step, window, vsg_count(f.step, window), u, v, src, exx, exx + NN, exx + 2 * NN,
                          st_));
            ck(hipMemcpyAsync(pin_out_.p, exx, sizeof(float) * 3 * NN, hipMemcpyDeviceToHost, st_));
            ck(hipMemcpyAsync(pin_ok_out_.p, src, NN, hipMemcpyDeviceToHost, st_));
            ck(hipMemcpyAsync(pin_counts_.p + 16, pass_fail, sizeof(int32_t) * kOutlierMaxPasses,
                              hipMemcpyDeviceToHost, st_));
        });
        ck(hipStreamSynchronize(st_));
        const float* hs = pin_out_.p;
        const uint8_t* hok = pin_ok_out_.p;
        int rej = 0;
        for (int pass = 0; pass < kOutlierMaxPasses; ++pass) rej += pin_counts_.p[16 + pass];
        int valid = 0, post = 0;
        bool trunc = false;
        for (int y = 0; y < gh && !trunc; ++y)
            for (int x = 0; x < gw; ++x) {
                const size_t i = (size_t)y * gw + x;
                if (!hok[i]) continue;
                if ((valid + 1) * 8 > out_capacity) {
                    trunc = true;
                    break;
                }
                if (hs[i] <= kStrainFail) {
                    post++;
                    continue;
                }
                float* o = out + (size_t)valid * 8;
                o[0] = (float)(f.rect_x + x * f.step);
                o[1] = (float)(f.rect_y + y * f.step);
                o[2] = f.u[i];
                o[3] = f.v[i];
                o[4] = hs[i];
                o[5] = hs[NN + i];
                o[6] = hs[2 * NN + i];
                o[7] = f.corr[i];
                valid++;
            }
        if (dropped_post) *dropped_post = post;
        if (dropped_outlier) *dropped_outlier = rej;
        return valid;
    } catch (const HipFail&) {
        return kErrInit;
    }
}

int validate_params(Params& p, int width, int height) { return validate(p, width, height); }
int vsg_window_node_count(int step, int window) { return vsg_count(step, window); }

}  // namespace sdic
