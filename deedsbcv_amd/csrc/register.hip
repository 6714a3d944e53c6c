// The symmetric multi-level registration driver (reference: deedsBCV0.cpp:139-356).
//
// Everything stays resident in HBM; the host only builds the MST trees.  The
// trees depend on nothing but the fixed image, the affine-warped moving image
// and the grid step (primsGraph is called on im1b and warped0 at every level,
// deedsBCV0.cpp:275 and :302), so all of them are started at the beginning:
// the GPU computes every level's edge SADs, host threads turn them into trees
// while the GPU runs MIND and the data costs.  Within a level the forward and
// backward directions are independent until consistentMappingCL (:311) and
// run on two streams, BP included.
//
// A registration is a PairRun: its phases (begin, per level pre-BP work, tree
// staging, post-BP work, outputs) are what the drivers interleave.  The
// single-pair driver runs one PairRun.  The multi-atlas driver runs many
// level-synchronously: every pair's pre-BP work of a level, then per direction
// the BP of all the pairs' trees in one launch sequence (one launch per round
// for all trees -- BP rounds are latency-bound, so 16 trees cost about as
// much as one), then every pair's post-BP work.  Every kernel sees the same inputs as in a
// registration of the pair alone, so the outputs are bit-identical.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <future>
#include <memory>
#include <mutex>
#include <string>
#include <thread>

#include "bp_host.h"
#include "common.h"
#include "kernels.h"
#include "mst_host.h"
#include "stats.h"

namespace deeds {

namespace {

__global__ void k_offset(float* __restrict__ out, const float* __restrict__ in, float add, long long n) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) out[p] = in[p] - add;  // in - threshold, deedsBCV0.cpp:139-144 / :352-354
}

void launch_offset(float* out, const float* in, float threshold, long long n, hipStream_t s) {
    if (n == 0) return;
    k_offset<<<ceil_div(n, 256), 256, 0, s>>>(out, in, threshold, n);
    DEEDS_LAUNCH_CHECK();
}

__global__ void k_zero(float* __restrict__ p, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

void launch_zero(float* p, long long n, hipStream_t s) {
    if (n == 0) return;
    k_zero<<<ceil_div(n, 256), 256, 0, s>>>(p, n);
    DEEDS_LAUNCH_CHECK();
}

struct Event {
    hipEvent_t e = nullptr;
    Event() { DEEDS_HIP(hipEventCreate(&e)); }
    ~Event() {
        if (e) (void)hipEventDestroy(e);
    }
    Event(const Event&) = delete;
    Event& operator=(const Event&) = delete;
};

double elapsed_s(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    DEEDS_HIP(hipEventElapsedTime(&ms, a, b));
    return ms * 1e-3;
}

// Host threads running tasks in submission order.  The destructor runs what
// is still queued and joins.
class TaskPool {
   public:
    explicit TaskPool(int threads) {
        for (int i = 0; i < std::max(1, threads); i++) workers_.emplace_back([this] { loop(); });
    }
    ~TaskPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& w : workers_) w.join();
    }
    TaskPool(const TaskPool&) = delete;
    TaskPool& operator=(const TaskPool&) = delete;

    template <class F>
    auto submit(F f) -> std::future<decltype(f())> {
        using R = decltype(f());
        auto task = std::make_shared<std::packaged_task<R()>>(std::move(f));
        std::future<R> fut = task->get_future();
        {
            std::lock_guard<std::mutex> lk(mu_);
            q_.emplace_back([task] { (*task)(); });
        }
        cv_.notify_one();
        return fut;
    }

   private:
    void loop() {
        for (;;) {
            std::function<void()> job;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [this] { return stop_ || !q_.empty(); });
                if (q_.empty()) return;
                job = std::move(q_.front());
                q_.pop_front();
            }
            job();
        }
    }
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<std::function<void()>> q_;
    bool stop_ = false;
    std::vector<std::thread> workers_;
};

struct LevelCfg {
    int step, hw, quant, len, len3, m1, n1, o1, sz1, mind_qs;
    bool new_mind;
    CostPlan plan;
};

std::vector<LevelCfg> level_cfgs(const deeds_schedule& S, int m, int n, int o) {
    std::vector<LevelCfg> lv(S.levels);
    for (int l = 0; l < S.levels; l++) {
        LevelCfg& c = lv[l];
        c.step = S.grid[l];
        c.hw = S.radius[l];
        c.quant = S.quant[l];
        c.len = 2 * c.hw + 1;
        c.len3 = c.len * c.len * c.len;
        c.m1 = m / c.step;
        c.n1 = n / c.step;
        c.o1 = o / c.step;
        c.sz1 = c.m1 * c.n1 * c.o1;
        c.mind_qs = (int)std::floor(0.5f * (float)c.quant + 1.0f);  // deedsBCV0.cpp:185
        c.new_mind = (l == 0) || (lv[l - 1].mind_qs != c.mind_qs);
        c.plan = datacost_plan(m, n, o, c.step, c.hw, (float)c.quant, S.alpha, 1);
    }
    return lv;
}

// Device scratch that pairs whose work is ordered on the same stream share:
// per direction the MIND temporaries, the warped image, its codes and the
// padded codes the data cost reads.
struct Scratch {
    float* mw[2][2];
    float* warped[2];
    uint64_t* mind_w[2];
    uint64_t* mpad[2];
};

Scratch make_scratch(deeds_ctx* ctx, size_t sz, size_t pad) {
    Scratch sh;
    for (int d = 0; d < 2; d++) {
        const std::string k = std::to_string(d);
        for (int b = 0; b < 2; b++) sh.mw[d][b] = ctx->get<float>("r.mindtmp." + k + std::to_string(b), 6 * sz);
        sh.warped[d] = ctx->get<float>("r.warped." + k, sz);
        sh.mind_w[d] = ctx->get<uint64_t>("r.mindw." + k, sz);
        sh.mpad[d] = ctx->get<uint64_t>("r.mpad." + k, pad);
    }
    return sh;
}

size_t max_padded(const std::vector<LevelCfg>& lv) {
    size_t p = 0;
    for (const LevelCfg& c : lv) p = std::max(p, (size_t)c.plan.mp * c.plan.np * c.plan.op);
    return p;
}

struct PairIn {
    const float* fixed;
    const float* moving;
    const float* X;
    int m, n, o;
    float* flow;
    float* warped;
    deeds_timing* timing;
};

// One symmetric registration, phase by phase.  Streams: ctx->stream (s0,
// forward direction and everything shared) and ctx->aux_stream (s1, the
// backward direction's pre-BP work and BP).  Workspace: ctx buffers named with
// `tag`, plus the shared scratch.
class PairRun {
   public:
    PairRun(deeds_ctx* ctx, const std::string& tag, const PairIn& in, const deeds_schedule& S, const Scratch& sh)
        : ctx_(ctx), tag_(tag), in_(in), S_(S), sh_(sh), L_(S.levels), s0_(ctx->stream), s1_(ctx->aux_stream) {
        const int m = in.m, n = in.n, o = in.o;
        sz_ = (long long)m * n * o;
        lv_ = level_cfgs(S, m, n, o);
        int max_qs = 0;
        size_t max_cost = 0, max_sz1 = 0;
        for (const LevelCfg& c : lv_) {
            max_qs = std::max(max_qs, c.mind_qs);
            max_cost = std::max(max_cost, (size_t)c.sz1 * c.len3);
            max_sz1 = std::max(max_sz1, (size_t)c.sz1);
        }
        for (const LevelCfg& c : lv_) hamming_ += 2.0 * (double)c.sz1 * c.len3 * (double)c.plan.samples;
        DEEDS_REQUIRE(m >= 2 * max_qs + 1 && n >= 2 * max_qs + 1 && o >= 2 * max_qs + 1,
                      "volume too small for the MIND patch radius of the schedule");
        DEEDS_REQUIRE(m <= mind_max_dim() && n <= mind_max_dim() && o <= mind_max_dim(),
                      "a dimension exceeds the box filter's LDS tile limit");
        const size_t sz = (size_t)sz_;
        im1_ = get<float>("im1", sz);
        im1b_ = get<float>("im1b", sz);
        warped0_ = get<float>("warped0", sz);
        mind_mov_ = get<uint64_t>("mind.mov", sz);
        mind_fix_ = get<uint64_t>("mind.fix", sz);
        for (int d = 0; d < 2; d++) {
            cost_[d] = get<float>("cost." + std::to_string(d), max_cost);
            inds_[d] = get<int16_t>("inds." + std::to_string(d), max_cost);
            // a pair registered alone (no tag) runs the single-tree BP, whose
            // selection walks composed tables; a batch's trees use the batched one
            comp_[d] = tag.empty() ? get<int16_t>("comp." + std::to_string(d), max_cost) : nullptr;
        }
        // control-grid fields: previous level (fwd u,v,w, bwd u,v,w), this level's u0 per direction, new fields
        const size_t sz2_0 = (size_t)(m / S.grid[0]) * (n / S.grid[0]) * (o / S.grid[0]);
        const size_t fmax = std::max(max_sz1, sz2_0);
        prevf_ = get<float>("prevf", 6 * fmax);
        newf_ = get<float>("newf", 6 * fmax);
        for (int d = 0; d < 2; d++) g0_[d] = get<float>("g0." + std::to_string(d), 3 * fmax);
        cmtmp_ = get<float>("cm", 6 * fmax);
        jac_ = get<double>("jac", 4 * (size_t)L_);
        ssd_ = get<double>("ssd", 2);
        red_part_ = get<double>("redpart", kRedPartDoubles);
        // MST inputs: edge SADs of every level on both images
        ec_off_.resize(L_);
        for (int l = 0; l < L_; l++) {
            ec_off_[l] = ec_total_;
            ec_total_ += 6 * (size_t)lv_[l].sz1;
        }
        d_ec_ = get<float>("ec", 2 * ec_total_);
        h_ec_ = static_cast<float*>(ctx->pinned(tag + "r.ec", 2 * ec_total_ * sizeof(float)));
        stats_bytes_ = (StatsScratch::bytes(sz_) + 255) & ~(size_t)255;
        d_stats_ = static_cast<char*>(ctx->buf(tag + "r.stats", 2 * stats_bytes_));
        h_stats_ = static_cast<char*>(ctx->pinned(tag + "r.stats", 2 * stats_bytes_));
        // pinned staging for every (level, direction) tree and a device area per direction
        stage_off_.resize(L_ + 1);
        size_t stage_max = 0;
        for (int l = 0; l < L_; l++) {
            const size_t b = (tree_stage_bytes(lv_[l].sz1) + 255) & ~(size_t)255;
            stage_off_[l] = stage_total_;
            stage_total_ += 2 * b;
            stage_max = std::max(stage_max, b);
        }
        stage_off_[L_] = stage_total_;
        h_stage_ = static_cast<char*>(ctx->pinned(tag + "r.stage", stage_total_));
        for (int d = 0; d < 2; d++) d_stage_[d] = ctx->buf(tag + "r.tree." + std::to_string(d), stage_max);
        for (int i = 0; i < 6 * L_; i++) ev_lvl_.emplace_back(new Event());
        for (int i = 0; i < 4 * L_; i++) ev_dc_.emplace_back(new Event());
        for (int i = 0; i < 2 * L_; i++) ev_pre_.emplace_back(new Event());
        for (int i = 0; i < L_; i++) ev_sad_.emplace_back(new Event());
        build_time_.assign(2 * L_, 0.0);
        build_start_.assign(2 * L_, 0.0);
        build_shape_.assign(2 * L_, 0.0);
        wait_level_.assign(L_, 0.0);
        trees_[0].resize(L_);
        trees_[1].resize(L_);
        for (int d = 0; d < 2; d++) {
            got_[d].resize(L_);
            have_[d].assign(L_, 0);
        }
    }

    int levels() const { return L_; }
    const LevelCfg& level(int l) const { return lv_[l]; }
    // Graph mode (a batch level's pre- and post-BP work captured into, or
    // replayed from, a HIP graph): the per-stage timing events are not recorded
    // (a captured event record is a dependency, not a timestamp), and the
    // cross-stream joins are made by the driver around the graph launch.
    void set_graph_mode(bool on) { graph_ = on; }
    // While a level is being captured, no other thread may touch the HIP
    // runtime on this context's behalf (an event query or a copy against a
    // capturing stream would invalidate the capture).  In that mode the tree
    // tasks wait for their edge SADs on a host flag, which wait_helpers sets
    // after this thread has waited for the SADs and the image statistics.
    void set_host_sad_wait(bool on) { host_sads_ = on; }
    void wait_helpers() {
        for (int l = 0; l < L_; l++) DEEDS_HIP(hipEventSynchronize(ev_sad_[l]->e));
        release_helpers();
        for (int d = 0; d < 2; d++) sd_fut_[d].wait();
    }
    // Sets the flag if it is not set yet (also on an error path, so that no
    // tree task waits for it forever).
    void release_helpers() {
        if (sad_set_) return;
        for (int l = 0; l < L_; l++) (void)hipEventSynchronize(ev_sad_[l]->e);
        sad_set_ = true;
        sad_promise_.set_value();
    }

    // Inputs, affine warp, image statistics and edge SADs (s0).
    void begin() {
        t_begin_ = std::chrono::steady_clock::now();
        DEEDS_HIP(hipEventRecord(ev_start_.e, s0_));
        // +1024 HU offset (deedsBCV0.cpp:139-144), affine warp (:210)
        const float identity[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
        A_ = affine_from_X(in_.X ? in_.X : identity);
        launch_offset(im1b_, in_.fixed, -1024.0f, sz_, s0_);
        launch_offset(im1_, in_.moving, -1024.0f, sz_, s0_);
        {
            // warpAffine with a zero field == the grid-fused warp from a zero 1x1x1 grid
            float* zero3 = g0_[0];
            launch_zero(zero3, 3, s0_);
            launch_warp_affine_grid(warped0_, im1_, A_, zero3, zero3 + 1, zero3 + 2, in_.m, in_.n, in_.o, 1, 1, 1,
                                    s0_);
        }
        // image statistics of both images (exact; chunks decomposed on the side streams)
        DEEDS_HIP(hipEventRecord(ev_w0_.e, s0_));
        hipEvent_t w0_event = ev_w0_.e;
        auto stats_task = [this, w0_event](int dir) {
            auto ts = std::chrono::steady_clock::now();
            hipStream_t ss = ctx_->side_stream[dir];
            DEEDS_HIP(hipSetDevice(ctx_->device));
            DEEDS_HIP(hipStreamWaitEvent(ss, w0_event, 0));
            StatsScratch sc;
            sc.dev = d_stats_ + dir * stats_bytes_;
            sc.host = h_stats_ + dir * stats_bytes_;
            float mean = 0, sd = 0;
            image_stats_device(dir ? warped0_ : im1b_, sz_, sc, ss, &mean, &sd);
            stats_time_[dir] = std::chrono::duration<double>(std::chrono::steady_clock::now() - ts).count();
            return sd;
        };
        sd_fut_[0] = std::async(std::launch::async, stats_task, 0).share();
        sd_fut_[1] = std::async(std::launch::async, stats_task, 1).share();
        // Edge SADs: the finest level first (its trees take longest to build),
        // then the first level (its trees are needed first), both on s0; the
        // levels between on s1, which has nothing else to do until the first
        // level's descriptors are done, so s0 goes on to them right away.
        auto sads = [&](int l, hipStream_t s) {
            const size_t cnt = 6 * (size_t)lv_[l].sz1;
            for (int dir = 0; dir < 2; dir++) {
                float* d = d_ec_ + dir * ec_total_ + ec_off_[l];
                launch_edge_sad(d, dir ? warped0_ : im1b_, lv_[l].step, in_.m, in_.n, in_.o, s);
                d2h(h_ec_ + dir * ec_total_ + ec_off_[l], d, cnt, s);
            }
            DEEDS_HIP(hipEventRecord(ev_sad_[l]->e, s));
        };
        sads(L_ - 1, s0_);
        if (L_ > 1) sads(0, s0_);
        if (L_ > 2) {
            DEEDS_HIP(hipStreamWaitEvent(s1_, w0_event, 0));
            for (int l = 1; l < L_ - 1; l++) sads(l, s1_);
        }
        // fields start at zero on the level-0 grid
        m2_ = in_.m / S_.grid[0];
        n2_ = in_.n / S_.grid[0];
        o2_ = in_.o / S_.grid[0];
        sz2_ = (size_t)m2_ * n2_ * o2_;
        launch_zero(prevf_, 6 * (long long)sz2_, s0_);
        DEEDS_HIP(hipEventRecord(ev_init_.e, s0_));
        DEEDS_HIP(hipStreamWaitEvent(s1_, ev_init_.e, 0));
    }

    // Queues the two trees of level l on the pool.  Prim's on the edge costs
    // runs before the image statistics are known; the exact weights then
    // confirm the order (else the tree is rebuilt from the weights).  For the
    // first level (needed first) and the finest (the longest build, which
    // the last BP waits for) the confirmation runs on a thread of its own
    // beside Prim's; the other levels confirm after Prim's on the building
    // thread, so fewer threads compete for the cores while the finest
    // level's Prim's runs (measured alone on the host: its tree 3-4 ms sooner).
    void submit_trees(TaskPool& pool, int l) {
        const bool check_beside = l == 0 || l == L_ - 1;
        for (int dir = 0; dir < 2; dir++) {
            const float* ec = h_ec_ + dir * ec_total_ + ec_off_[l];
            const LevelCfg c = lv_[l];
            std::shared_future<float> sdf = sd_fut_[dir];
            double* spent = &build_time_[2 * l + dir];
            double* began = &build_start_[2 * l + dir];
            double* shaped = &build_shape_[2 * l + dir];
            const auto t_begin = t_begin_;
            hipEvent_t sad_event = ev_sad_[l]->e;
            const bool host_wait = host_sads_;
            std::shared_future<void> sads = sad_ready_;
            // the schedule is packed into this (level, direction)'s pinned slot by the
            // building thread, so the launching thread only enqueues the copy
            char* slot = h_stage_ + stage_off_[l] + dir * ((stage_off_[l + 1] - stage_off_[l]) / 2);
            void* d_stage = d_stage_[dir];
            trees_[dir][l] = pool.submit([ec, c, sdf, spent, began, shaped, t_begin, sad_event, slot, d_stage,
                                          host_wait, sads, check_beside]() {
                if (host_wait)
                    sads.wait();
                else
                    DEEDS_HIP(hipEventSynchronize(sad_event));
                auto tb = std::chrono::steady_clock::now();
                *began = std::chrono::duration<double>(tb - t_begin).count();
                auto confirm = [ec, c, sdf]() {
                    return weights_distinguish_costs(ec, c.step, c.m1, c.n1, c.o1, sdf.get());
                };
                std::future<bool> check;
                if (check_beside) check = std::async(std::launch::async, confirm);
                std::unique_ptr<Tree> t(new Tree());
                // (declined for non-integer SADs: then Prim's runs once, on the weights)
                const bool by_cost = build_tree_shape(ec, c.step, c.m1, c.n1, c.o1, *t);
                *shaped = std::chrono::duration<double>(std::chrono::steady_clock::now() - tb).count();
                const bool same_order = by_cost && (check_beside ? check.get() : confirm());
                const float sd = sdf.get();
                if (same_order)
                    finish_tree(ec, sd, c.step, *t);
                else
                    build_tree(ec, sd, c.step, c.m1, c.n1, c.o1, *t);
                StagedTree st;
                st.dt = pack_tree(*t, slot, d_stage, &st.bytes);
                st.tree = std::move(t);
                *spent = std::chrono::duration<double>(std::chrono::steady_clock::now() - tb).count();
                return st;
            });
        }
    }

    // Node steps on the bottom-up pass's critical path as the rounds run it
    // (each round costs its longest segment), or -1 while the tree is not built.
    long long bp_steps(int l, int dir) {
        if (!have_[dir][l]) {
            if (trees_[dir][l].wait_for(std::chrono::seconds(0)) != std::future_status::ready) return -1;
            got_[dir][l] = trees_[dir][l].get();
            have_[dir][l] = true;
        }
        return got_[dir][l].tree->bp_steps;
    }

    // The direction whose tree of level l is built first (the forward one
    // when both are, or once both are); the host's wait counts as tree wait.
    int first_built(int l) {
        auto tw = std::chrono::steady_clock::now();
        int dir = 0;
        while (true) {
            if (bp_steps(l, 0) >= 0) break;
            if (bp_steps(l, 1) >= 0) {
                dir = 1;
                break;
            }
            std::this_thread::sleep_for(std::chrono::microseconds(20));
        }
        const double waited = std::chrono::duration<double>(std::chrono::steady_clock::now() - tw).count();
        wait_time_ += waited;
        wait_level_[l] += waited;
        return dir;
    }

    // The direction whose data cost runs first, alone on the GPU (the other
    // waits for it): its BP then starts half a data-cost phase earlier while
    // the total data-cost time stays the same.  When both trees of level l are
    // built, the one whose BP is expected to take longer; otherwise the
    // forward direction.
    int first_direction(int l) {
        const long long a = bp_steps(l, 0), b = bp_steps(l, 1);
        if (a < 0 || b < 0) return 0;
        return a >= b ? 0 : 1;
    }

    // The MIND codes of warped0 and im1b depend on the patch radius only, so
    // the next level that changes the radius can take them from a second pair
    // of buffers filled beforehand.  Queued on s0 after level l's pre-BP
    // work, they run while the GPU would otherwise wait for level l's trees
    // (the first level's trees are built while the GPU has little else to
    // do).  The temporaries are the forward direction's, used on s0 only.
    void mind_ahead(int l) {
        if (graph_ || ahead_level_ >= 0) return;
        int l2 = l + 1;
        while (l2 < L_ && !lv_[l2].new_mind) l2++;
        if (l2 >= L_) return;
        const size_t sz = (size_t)in_.m * in_.n * in_.o;
        if (!mind_mov2_) {
            mind_mov2_ = get<uint64_t>("mind.mov2", sz);
            mind_fix2_ = get<uint64_t>("mind.fix2", sz);
        }
        const int qs = lv_[l2].mind_qs;
        launch_descriptor(mind_mov2_, warped0_, in_.m, in_.n, in_.o, qs, sh_.mw[0][0], sh_.mw[0][1], s0_);
        launch_descriptor(mind_fix2_, im1b_, in_.m, in_.n, in_.o, qs, sh_.mw[0][0], sh_.mw[0][1], s0_);
        ahead_level_ = l2;
    }

    // Pre-BP work of both directions: MIND of the images when the patch radius
    // changes, then per direction upsample, warp, MIND, padding, data cost
    // (the `first` direction's data cost before the other's; -1: concurrent).
    void pre_bp(int l, int first = -1) {
        const LevelCfg& c = lv_[l];
        if (!graph_) DEEDS_HIP(hipEventRecord(EV(l, 0), s0_));
        if (c.new_mind && ahead_level_ == l) {
            // computed ahead (mind_ahead): the codes of this patch radius take the place of the previous ones
            std::swap(mind_mov_, mind_mov2_);
            std::swap(mind_fix_, mind_fix2_);
            ahead_level_ = -1;
        } else if (c.new_mind) {
            // both directions' MIND of warped0 / im1b (deedsBCV0.cpp:230-236)
            launch_descriptor(mind_mov_, warped0_, in_.m, in_.n, in_.o, c.mind_qs, sh_.mw[0][0], sh_.mw[0][1], s0_);
            launch_descriptor(mind_fix_, im1b_, in_.m, in_.n, in_.o, c.mind_qs, sh_.mw[0][0], sh_.mw[0][1], s0_);
        }
        DEEDS_HIP(hipEventRecord(ev_mind_.e, s0_));
        DEEDS_HIP(hipStreamWaitEvent(s1_, ev_mind_.e, 0));
        for (int di = 0; di < 2; di++) {
            const int dir = first == 1 ? 1 - di : di;  // the first direction's work is queued first
            hipStream_t s = dir ? s1_ : s0_;
            const float* pf = prevf_ + dir * 3 * sz2_;
            float* u0 = g0_[dir];
            const size_t n1 = c.sz1;
            // upsampleDeformationsCL to this grid (:254 / :283)
            launch_upsample(u0, u0 + n1, u0 + 2 * n1, pf, pf + sz2_, pf + 2 * sz2_, c.m1, c.n1, c.o1, m2_, n2_, o2_,
                            s);
            // upsample to full resolution fused with the warp (:255-257 / :284-285)
            if (dir == 0)
                launch_warp_affine_grid(sh_.warped[0], im1_, A_, u0, u0 + n1, u0 + 2 * n1, in_.m, in_.n, in_.o, c.m1,
                                        c.n1, c.o1, s);
            else
                launch_warp_image_grid(sh_.warped[1], im1b_, u0, u0 + n1, u0 + 2 * n1, in_.m, in_.n, in_.o, c.m1,
                                       c.n1, c.o1, s);
            if (dir == 0 && !graph_) DEEDS_HIP(hipEventRecord(EV(l, 1), s));
            launch_descriptor(sh_.mind_w[dir], sh_.warped[dir], in_.m, in_.n, in_.o, c.mind_qs, sh_.mw[dir][0],
                              sh_.mw[dir][1], s);
            if (dir == 0 && !graph_) DEEDS_HIP(hipEventRecord(EV(l, 2), s));
            launch_pad_codes(sh_.mpad[dir], sh_.mind_w[dir], in_.m, in_.n, in_.o, c.plan.pad, s);
            // the second direction's data cost waits for the first's
            if (first >= 0 && dir != first) DEEDS_HIP(hipStreamWaitEvent(s, DC(l, first, 1), 0));
            if (!graph_) DEEDS_HIP(hipEventRecord(DC(l, dir, 0), s));
            launch_datacost(cost_[dir], dir ? mind_mov_ : mind_fix_, sh_.mpad[dir], in_.m, in_.n, in_.o, c.step,
                            c.hw, c.quant, c.plan, s);
            if (!graph_) DEEDS_HIP(hipEventRecord(DC(l, dir, 1), s));
            if (dir == 0 && !graph_) DEEDS_HIP(hipEventRecord(EV(l, 3), s));
        }
    }

    // Waits for the level-l tree of direction `dir`, stages it on that
    // direction's stream and appends its BP item (regularisationCL, :276 / :303).
    void bp_item(int l, int dir, std::vector<BPBatchItem>& items) {
        const LevelCfg& c = lv_[l];
        hipStream_t s = dir ? s1_ : s0_;
        const size_t half = (stage_off_[l + 1] - stage_off_[l]) / 2;
        auto tw = std::chrono::steady_clock::now();
        StagedTree st = have_[dir][l] ? std::move(got_[dir][l]) : trees_[dir][l].get();
        have_[dir][l] = false;
        const double waited = std::chrono::duration<double>(std::chrono::steady_clock::now() - tw).count();
        wait_time_ += waited;
        wait_level_[l] += waited;
        DEEDS_HIP(hipEventRecord(ev_pre_[2 * l + dir]->e, s));
        char* slot = h_stage_ + stage_off_[l] + dir * half;
        DEEDS_HIP(hipMemcpyAsync(d_stage_[dir], slot, st.bytes, hipMemcpyHostToDevice, s));
        tree_[dir] = std::move(st.tree);
        const BPDevTree dt = st.dt;
        float* u0 = g0_[dir];
        const size_t n1 = c.sz1;
        float* nf = newf_ + dir * 3 * n1;
        items.push_back(BPBatchItem{tree_[dir].get(), dt, cost_[dir], inds_[dir], comp_[dir], u0, u0 + n1,
                                    u0 + 2 * n1, nf, nf + n1, nf + 2 * n1});
    }

    // Marks the end of the forward direction's BP of level l on s0.
    void fwd_bp_done(int l) { DEEDS_HIP(hipEventRecord(EV(l, 4), s0_)); }

    // After both directions' BP of level l (s0 waits for s1):
    // consistentMappingCL (:311) on the new fields, which then become "previous".
    void post_bp(int l) {
        const LevelCfg& c = lv_[l];
        const size_t n1 = c.sz1;
        float* f[6] = {newf_, newf_ + n1, newf_ + 2 * n1, newf_ + 3 * n1, newf_ + 4 * n1, newf_ + 5 * n1};
        launch_consistent(f, cmtmp_, c.m1, c.n1, c.o1, c.step, s0_);
        launch_jacobian(jac_ + 4 * l, red_part_, newf_, newf_ + n1, newf_ + 2 * n1, c.m1, c.n1, c.o1, c.step, s0_);
        DEEDS_HIP(hipMemcpyAsync(prevf_, newf_, 6 * n1 * sizeof(float), hipMemcpyDeviceToDevice, s0_));
        if (!graph_) {
            DEEDS_HIP(hipEventRecord(EV(l, 5), s0_));
            DEEDS_HIP(hipStreamWaitEvent(s1_, EV(l, 5), 0));
        }
        post_bp_state(l);
    }
    // The host state post_bp leaves for the next level (also when its work
    // came from a replayed graph): the previous grid is this level's.
    void post_bp_state(int l) {
        const LevelCfg& c = lv_[l];
        m2_ = c.m1;
        n2_ = c.n1;
        o2_ = c.o1;
        sz2_ = c.sz1;
    }

    // Outputs (deedsBCV0.cpp:339-356), on s0.
    void finish() {
        DEEDS_HIP(hipEventRecord(ev_stop_.e, s0_));
        DEEDS_HIP(hipMemcpyAsync(in_.flow, prevf_, 3 * sz2_ * sizeof(float), hipMemcpyDeviceToDevice, s0_));
        if (in_.warped) {
            float* wtmp = sh_.warped[0];
            launch_warp_affine_grid(wtmp, im1_, A_, prevf_, prevf_ + sz2_, prevf_ + 2 * sz2_, in_.m, in_.n, in_.o,
                                    m2_, n2_, o2_, s0_);
            launch_ssd(ssd_, red_part_, im1b_, wtmp, im1_, sz_, s0_);
            launch_offset(in_.warped, wtmp, 1024.0f, sz_, s0_);
        }
    }

    // After both streams have drained: the timing record.
    void report() {
        deeds_timing* timing = in_.timing;
        if (!timing) return;
        std::memset(timing, 0, sizeof(*timing));
        timing->levels = L_;
        for (int l = 0; l < L_ && !graph_; l++) {
            timing->transf[l] = elapsed_s(EV(l, 0), EV(l, 1)) + elapsed_s(EV(l, 4), EV(l, 5));
            timing->mind[l] = elapsed_s(EV(l, 1), EV(l, 2));
            timing->data[l] = elapsed_s(EV(l, 2), EV(l, 3));
            timing->mst_reg[l] = elapsed_s(EV(l, 3), EV(l, 4));
        }
        timing->total = elapsed_s(ev_start_.e, ev_stop_.e);
        timing->hamming_evals = hamming_;
        timing->tree_wait_s = wait_time_;
        timing->stats_s = std::max(stats_time_[0], stats_time_[1]);
        for (int l = 0; l < L_; l++) {
            double lvl = 0.0;
            if (!graph_)
                for (int dir = 0; dir < 2; dir++) lvl += elapsed_s(DC(l, dir, 0), DC(l, dir, 1));
            timing->datacost_level_s[l] = lvl;
            timing->datacost_level_evals[l] = 2.0 * lv_[l].sz1 * lv_[l].len3 * (double)lv_[l].plan.samples;
            timing->datacost_kernel_s += lvl;
            timing->tree_build_s += build_time_[2 * l] + build_time_[2 * l + 1];
            timing->tree_build_level_s[l] = std::max(build_time_[2 * l], build_time_[2 * l + 1]);
            timing->tree_ready_level_s[l] = std::max(build_start_[2 * l] + build_time_[2 * l],
                                                     build_start_[2 * l + 1] + build_time_[2 * l + 1]);
            timing->tree_shape_level_s[l] = std::max(build_shape_[2 * l], build_shape_[2 * l + 1]);
            timing->tree_wait_level_s[l] = wait_level_[l];
            timing->datacost_launches += graph_ ? 0 : 2;
            for (int dir = 0; dir < 2 && !graph_; dir++)
                timing->bp_idle_level_s[l] += std::max(0.0, elapsed_s(DC(l, dir, 1), ev_pre_[2 * l + dir]->e));
            timing->bp_level_s[l] = elapsed_s(ev_pre_[2 * l]->e, EV(l, 4));
        }
        std::vector<double> hj(4 * (size_t)L_);
        DEEDS_HIP(hipMemcpy(hj.data(), jac_, hj.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int l = 0; l < L_; l++) {
            const double N = (double)lv_[l].sz1;
            const double mean = hj[4 * l] / N;
            const double var = N > 1 ? (hj[4 * l + 1] - N * mean * mean) / (N - 1) : 0.0;
            timing->jac_std[l] = (float)std::sqrt(std::max(var, 0.0));
            timing->jac_neg[l] = (float)(hj[4 * l + 2] / N);
        }
        if (in_.warped) {
            double hs[2];
            DEEDS_HIP(hipMemcpy(hs, ssd_, 2 * sizeof(double), hipMemcpyDeviceToHost));
            timing->ssd_before = (float)(hs[0] / (double)sz_);
            timing->ssd_after = (float)(hs[1] / (double)sz_);
        }
    }

   private:
    template <class T>
    T* get(const std::string& name, size_t count) {
        return ctx_->get<T>(tag_ + "r." + name, count);
    }
    hipEvent_t EV(int l, int k) { return ev_lvl_[6 * l + k]->e; }
    hipEvent_t DC(int l, int dir, int end) { return ev_dc_[4 * l + 2 * dir + end]->e; }

    deeds_ctx* ctx_;
    std::string tag_;
    PairIn in_;
    const deeds_schedule& S_;
    Scratch sh_;
    int L_;
    hipStream_t s0_, s1_;
    long long sz_ = 0;
    std::vector<LevelCfg> lv_;
    Affine A_{};
    float *im1_, *im1b_, *warped0_;
    uint64_t *mind_mov_, *mind_fix_;
    uint64_t *mind_mov2_ = nullptr, *mind_fix2_ = nullptr;  // codes computed ahead (mind_ahead)
    int ahead_level_ = -1;
    float* cost_[2];
    int16_t* inds_[2];
    int16_t* comp_[2];
    float *prevf_, *newf_, *cmtmp_;
    float* g0_[2];
    double *jac_, *ssd_, *red_part_;
    size_t ec_total_ = 0;
    std::vector<size_t> ec_off_;
    float* d_ec_;
    float* h_ec_;
    size_t stats_bytes_ = 0;
    char* d_stats_;
    char* h_stats_;
    std::vector<size_t> stage_off_;
    size_t stage_total_ = 0;
    char* h_stage_;
    void* d_stage_[2];
    int m2_ = 0, n2_ = 0, o2_ = 0;
    size_t sz2_ = 0;
    double hamming_ = 0.0;
    bool graph_ = false;
    bool host_sads_ = false;
    bool sad_set_ = false;
    std::promise<void> sad_promise_;
    std::shared_future<void> sad_ready_ = sad_promise_.get_future().share();
    Event ev_start_, ev_stop_, ev_init_, ev_mind_, ev_w0_;
    std::vector<std::unique_ptr<Event>> ev_sad_;  // per level: its edge SADs are on the host
    std::vector<std::unique_ptr<Event>> ev_lvl_, ev_dc_, ev_pre_;
    double stats_time_[2] = {0.0, 0.0};
    std::shared_future<float> sd_fut_[2];
    std::vector<double> build_time_, build_start_, build_shape_;
    std::chrono::steady_clock::time_point t_begin_;
    double wait_time_ = 0.0;
    std::vector<double> wait_level_;
    // a level's tree with its schedule already packed into the pinned slot
    struct StagedTree {
        std::unique_ptr<Tree> tree;
        BPDevTree dt{};
        size_t bytes = 0;
    };
    std::vector<std::future<StagedTree>> trees_[2];
    std::vector<StagedTree> got_[2];  // trees taken early by bp_steps
    std::vector<char> have_[2];
    std::unique_ptr<Tree> tree_[2];
};

// HIP graphs of a batch's levels.  A group of pairs runs each level's pre-BP
// work (MIND, warps, padding, data costs of every pair) and post-BP work
// (consistent mapping, logging Jacobian, field copy) as fixed launch
// sequences whose shapes and buffers depend only on the pairs and the
// schedule, so they are captured into one graph per level and phase and
// replayed when the same group comes again (the multi-atlas case: one target,
// the same atlases, call after call).  The BP between them depends on the
// host-built trees and stays on the streams.  The first call with a given key
// (pairs' buffers, sizes and affines, the schedule, the context's workspace
// addresses) runs eagerly, the second captures and launches, later ones
// replay.  DEEDS_GRAPHS=0 turns graphs off.
std::atomic<long long> g_graph_captures{0}, g_graph_replays{0};

struct LevelGraphs {
    std::vector<uint64_t> key, pending;
    std::vector<hipGraphExec_t> pre, post;
    void clear() {
        for (hipGraphExec_t g : pre)
            if (g) (void)hipGraphExecDestroy(g);
        for (hipGraphExec_t g : post)
            if (g) (void)hipGraphExecDestroy(g);
        pre.clear();
        post.clear();
        key.clear();
    }
    ~LevelGraphs() { clear(); }
};

// -1: not set yet (the environment decides), else 0 / 1 from deeds_set_batch_graphs
std::atomic<int> g_graphs_on{-1};

bool graphs_enabled() {
    int v = g_graphs_on.load();
    if (v < 0) {
        const char* e = getenv("DEEDS_GRAPHS");
        v = (e && e[0] == '0') ? 0 : 1;
        int expect = -1;
        g_graphs_on.compare_exchange_strong(expect, v);
        v = g_graphs_on.load();
    }
    return v != 0;
}

std::vector<uint64_t> group_key(deeds_ctx* ctx, const std::vector<PairIn>& pairs, const deeds_schedule& S) {
    std::vector<uint64_t> k;
    auto put = [&](uint64_t v) { k.push_back(v); };
    put(pairs.size());
    put((uint64_t)S.levels);
    uint32_t ab;
    std::memcpy(&ab, &S.alpha, 4);
    put(ab);
    for (int l = 0; l < S.levels; l++) put(((uint64_t)S.grid[l] << 40) | ((uint64_t)S.radius[l] << 20) | S.quant[l]);
    for (const PairIn& p : pairs) {
        put((uintptr_t)p.fixed);
        put((uintptr_t)p.moving);
        put((uintptr_t)p.flow);
        put((uintptr_t)p.warped);
        put(((uint64_t)p.m << 42) | ((uint64_t)p.n << 21) | (uint64_t)p.o);
        for (int i = 0; i < 16; i++) {
            uint32_t b = 0;
            if (p.X) std::memcpy(&b, &p.X[i], 4);
            put(p.X ? b : 0xffffffffull);
        }
    }
    for (const auto& kv : ctx->ws) {
        put((uintptr_t)kv.second.ptr);
        put(kv.second.bytes);
    }
    return k;
}

// Captures what `enqueue` queues on s0 (and on streams that join it) into a
// graph and instantiates it; a failed capture is ended before the error
// propagates, so the stream leaves capture mode.
template <class F>
hipGraphExec_t capture_graph(hipStream_t s0, F&& enqueue) {
    DEEDS_HIP(hipStreamBeginCapture(s0, hipStreamCaptureModeRelaxed));
    try {
        enqueue();
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(s0, &g);
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        throw;
    }
    hipGraph_t g = nullptr;
    DEEDS_HIP(hipStreamEndCapture(s0, &g));
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) fail(DEEDS_ERR_HIP, std::string("hipGraphInstantiate failed: ") + hipGetErrorString(e));
    return ex;
}

// BP staging of one level for `trees` trees with `nodes` nodes in total,
// 256-byte aligned.
size_t level_bp_bytes(int trees, long long nodes) { return (bp_batch_bytes(trees, nodes) + 255) & ~(size_t)255; }

// Runs the pairs level-synchronously on ctx (see the file comment).  Pairs
// are validated by the callers.
void run_level_sync(deeds_ctx* ctx, const std::vector<PairIn>& pairs, const deeds_schedule& S, int tree_threads) {
    const int P = (int)pairs.size();
    if (P == 0) return;
    const int L = S.levels;
    hipStream_t s0 = ctx->stream, s1 = ctx->aux_stream;
    // shared scratch at the largest pair's size, allocated before any work is queued
    size_t max_sz = 0, max_pad = 0;
    std::vector<long long> level_nodes(L, 0);
    for (const PairIn& p : pairs) {
        const std::vector<LevelCfg> lv = level_cfgs(S, p.m, p.n, p.o);
        max_sz = std::max(max_sz, (size_t)p.m * p.n * p.o);
        max_pad = std::max(max_pad, max_padded(lv));
        for (int l = 0; l < L; l++) level_nodes[l] += lv[l].sz1;
    }
    const Scratch sh = make_scratch(ctx, max_sz, max_pad);
    // BP staging per (level, direction): pinned slots for the whole call, a
    // device area per direction reused level after level on its stream
    std::vector<size_t> bpq_off(2 * L + 1, 0);
    size_t bpq_max = 0;
    for (int l = 0; l < L; l++)
        for (int dir = 0; dir < 2; dir++) {
            const size_t b = level_bp_bytes(P, level_nodes[l]);
            bpq_off[2 * l + dir + 1] = bpq_off[2 * l + dir] + b;
            bpq_max = std::max(bpq_max, b);
        }
    char* h_bpq = static_cast<char*>(ctx->pinned("r.bpq", bpq_off[2 * L]));
    void* d_bpq[2] = {ctx->buf("r.bpq.0", bpq_max), ctx->buf("r.bpq.1", bpq_max)};
    std::vector<std::unique_ptr<PairRun>> runs;
    for (int p = 0; p < P; p++)
        runs.emplace_back(new PairRun(ctx, P == 1 ? std::string() : "b" + std::to_string(p) + ".", pairs[p], S, sh));
    // graphs for the group's pre- and post-BP work of every level
    LevelGraphs* G = nullptr;
    bool replay = false, capture = false;
    if (P > 1 && graphs_enabled()) {
        if (!ctx->graph_cache) ctx->graph_cache = std::make_shared<LevelGraphs>();
        G = static_cast<LevelGraphs*>(ctx->graph_cache.get());
        const std::vector<uint64_t> key = group_key(ctx, pairs, S);
        replay = G->key == key && (int)G->pre.size() == L && (int)G->post.size() == L;
        capture = !replay && G->pending == key;
        if (!replay) {
            G->clear();
            G->pending = key;
        }
        if (capture) {
            G->pre.assign(L, nullptr);
            G->post.assign(L, nullptr);
        }
        for (auto& r : runs) {
            r->set_graph_mode(replay || capture);
            r->set_host_sad_wait(capture);
        }
    }
    Event ev_graph;
    // after a graph launched on s0, s1 continues behind it
    auto s1_behind_s0 = [&] {
        DEEDS_HIP(hipEventRecord(ev_graph.e, s0));
        DEEDS_HIP(hipStreamWaitEvent(s1, ev_graph.e, 0));
    };
    // declared after the runs: destroyed (drained and joined) before them
    TaskPool pool(tree_threads);
    // destroyed before the pool: no tree task is left waiting for a host flag
    struct ReleaseHelpers {
        std::vector<std::unique_ptr<PairRun>>& runs;
        bool on;
        ~ReleaseHelpers() {
            if (on)
                for (auto& r : runs) r->release_helpers();
        }
    } release_guard{runs, capture};
    for (auto& r : runs) r->begin();
    // trees in the order they are needed: level by level over the pairs
    for (int l = 0; l < L; l++)
        for (auto& r : runs) r->submit_trees(pool, l);
    if (capture)
        for (auto& r : runs) r->wait_helpers();
    Event ev_join;
    for (int l = 0; l < L; l++) {
        // one pair: the direction with the longer expected BP computes its data cost first
        if (P == 1) {
            runs[0]->pre_bp(l, runs[0]->first_direction(l));
            if (l == 0) runs[0]->mind_ahead(0);
        } else if (replay) {
            DEEDS_HIP(hipGraphLaunch(G->pre[l], s0));
            s1_behind_s0();
        } else if (capture) {
            G->pre[l] = capture_graph(s0, [&] {
                for (auto& r : runs) r->pre_bp(l);
                // the backward direction's branch (s1) joins s0 again
                DEEDS_HIP(hipEventRecord(ev_graph.e, s1));
                DEEDS_HIP(hipStreamWaitEvent(s0, ev_graph.e, 0));
            });
            DEEDS_HIP(hipGraphLaunch(G->pre[l], s0));
            s1_behind_s0();
        } else {
            for (auto& r : runs) r->pre_bp(l);
        }
        // each direction's trees of all pairs in one BP launch sequence on that direction's stream
        const LevelCfg& c = runs[0]->level(l);
        // one pair: the BP whose tree is built first is launched first (the
        // two directions' BPs run on their own streams)
        const int d0 = P == 1 ? runs[0]->first_built(l) : 0;
        for (int k = 0; k < 2; k++) {
            const int dir = k ? 1 - d0 : d0;
            std::vector<BPBatchItem> items;
            for (auto& r : runs) r->bp_item(l, dir, items);
            const size_t off = bpq_off[2 * l + dir];
            run_bp_batch(items, c.hw, (float)c.quant, h_bpq + off, d_bpq[dir], bpq_off[2 * l + dir + 1] - off,
                         dir ? s1 : s0);
            if (dir == 0)
                for (auto& r : runs) r->fwd_bp_done(l);
        }
        DEEDS_HIP(hipEventRecord(ev_join.e, s1));
        DEEDS_HIP(hipStreamWaitEvent(s0, ev_join.e, 0));
        if (replay) {
            DEEDS_HIP(hipGraphLaunch(G->post[l], s0));
            for (auto& r : runs) r->post_bp_state(l);
            s1_behind_s0();
        } else if (capture) {
            G->post[l] = capture_graph(s0, [&] {
                for (auto& r : runs) r->post_bp(l);
            });
            DEEDS_HIP(hipGraphLaunch(G->post[l], s0));
            s1_behind_s0();
        } else {
            for (auto& r : runs) r->post_bp(l);
        }
    }
    if (G && capture) {
        G->key = G->pending;
        G->pending.clear();
        g_graph_captures++;
    } else if (G && replay) {
        g_graph_replays++;
    }
    for (auto& r : runs) r->finish();
    DEEDS_HIP(hipStreamSynchronize(s0));
    DEEDS_HIP(hipStreamSynchronize(s1));
    for (auto& r : runs) r->report();
}

}  // namespace

void register_pair(deeds_ctx* ctx, const float* d_fixed, const float* d_moving, const float* X, int m, int n, int o,
                   const deeds_schedule& S, float* d_flow, float* d_warped, deeds_timing* timing) {
    const std::vector<PairIn> one{PairIn{d_fixed, d_moving, X, m, n, o, d_flow, d_warped, timing}};
    run_level_sync(ctx, one, S, 2 * S.levels);
}

void set_batch_graphs(bool on) { g_graphs_on = on ? 1 : 0; }

std::string graph_report() {
    return "{\"batch_graph_captures\": " + std::to_string(g_graph_captures.load()) +
           ", \"batch_graph_replays\": " + std::to_string(g_graph_replays.load()) + "}";
}

void register_group(deeds_ctx* ctx, const deeds_pair* pairs, int count, const deeds_schedule& S,
                    deeds_timing* timings) {
    std::vector<PairIn> in;
    for (int i = 0; i < count; i++) {
        const deeds_pair& p = pairs[i];
        in.push_back(PairIn{p.fixed, p.moving, p.X, p.m, p.n, p.o, p.flow, p.warped, timings ? &timings[i] : nullptr});
    }
    const int hc = (int)std::thread::hardware_concurrency();
    run_level_sync(ctx, in, S, std::max(2, std::min(16, hc)));
}

}  // namespace deeds

using namespace deeds;

void deeds::check_schedule(const deeds_schedule* S, int m, int n, int o) {
    DEEDS_REQUIRE(S, "schedule is NULL");
    DEEDS_REQUIRE(S->levels >= 1 && S->levels <= DEEDS_MAX_LEVELS, "levels must be in [1, 10]");
    DEEDS_REQUIRE(m > 0 && n > 0 && o > 0 && (long long)m * n * o < (1LL << 31), "bad volume dimensions");
    for (int l = 0; l < S->levels; l++) {
        DEEDS_REQUIRE(S->grid[l] >= 1 && m / S->grid[l] >= 1 && n / S->grid[l] >= 1 && o / S->grid[l] >= 1,
                      "grid spacing must be >= 1 and not exceed a dimension");
        DEEDS_REQUIRE(S->radius[l] >= 0 && S->radius[l] <= kMaxBPRadius, "search radius must be in [0, 10]");
        DEEDS_REQUIRE(S->quant[l] >= 1, "quantisation must be >= 1");
    }
    DEEDS_REQUIRE(std::isfinite(S->alpha) && S->alpha > 0.0f, "alpha must be positive");
}


extern "C" {

int deeds_ctx_register(deeds_ctx* ctx, const float* d_fixed, const float* d_moving, const float* X, int m, int n,
                       int o, const deeds_schedule* sched, float* d_flow, float* d_warped, deeds_timing* timing) {
    if (!ctx) return DEEDS_ERR_ARG;
    return guarded(ctx, [&] {
        check_schedule(sched, m, n, o);
        DEEDS_REQUIRE(d_fixed && d_moving && d_flow, "null pointer");
        register_pair(ctx, d_fixed, d_moving, X, m, n, o, *sched, d_flow, d_warped, timing);
    });
}

int deeds_register(const float* fixed, const float* moving, const float* X, int m, int n, int o,
                   const deeds_schedule* sched, float* flow_out, float* warped_out, deeds_timing* timing) {
    return guarded(nullptr, [&] {
        DEEDS_REQUIRE(fixed && moving && flow_out, "null pointer");
        check_schedule(sched, m, n, o);
        deeds_ctx* c = default_ctx();
        DEEDS_HIP(hipSetDevice(c->device));
        size_t sz = (size_t)m * n * o;
        int g = sched->grid[sched->levels - 1];
        size_t sz1 = (size_t)(m / g) * (n / g) * (o / g);
        float* d_in = c->get<float>("h.reg.in", 2 * sz);
        float* d_out = c->get<float>("h.reg.out", sz + 3 * sz1);
        h2d(d_in, fixed, sz, c->stream);
        h2d(d_in + sz, moving, sz, c->stream);
        int rc = deeds_ctx_register(c, d_in, d_in + sz, X, m, n, o, sched, d_out + sz, warped_out ? d_out : nullptr,
                                    timing);
        if (rc) fail(rc, c->err);
        d2h(flow_out, d_out + sz, 3 * sz1, c->stream);
        if (warped_out) d2h(warped_out, d_out, sz, c->stream);
        DEEDS_HIP(hipStreamSynchronize(c->stream));
    });
}

int deeds_ctx_apply_seg(deeds_ctx* ctx, int16_t* d_out, const int16_t* d_in, const float* X, const float* d_flow,
                        int grid_step, int m, int n, int o) {
    if (!ctx) return DEEDS_ERR_ARG;
    return guarded(ctx, [&] {
        DEEDS_REQUIRE(d_out && d_in && d_flow, "null pointer");
        DEEDS_REQUIRE(d_out != d_in, "out must not alias in");
        DEEDS_REQUIRE(m > 0 && n > 0 && o > 0 && (long long)m * n * o < (1LL << 31), "bad volume dimensions");
        DEEDS_REQUIRE(grid_step >= 1 && m / grid_step >= 1 && n / grid_step >= 1 && o / grid_step >= 1,
                      "bad grid step");
        const size_t sz = (size_t)m * n * o;
        const int m1 = m / grid_step, n1 = n / grid_step, o1 = o / grid_step;
        const size_t sz1 = (size_t)m1 * n1 * o1;
        float* full = ctx->get<float>("apply.full", 3 * sz);
        // applyBCV.cpp:161-180: the full-resolution field, then the nearest-neighbour warp
        launch_upsample(full, full + sz, full + 2 * sz, d_flow, d_flow + sz1, d_flow + 2 * sz1, m, n, o, m1, n1, o1,
                        ctx->stream);
        const float identity[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
        const Affine A = affine_from_X(X ? X : identity);
        launch_warp_affine_s(d_out, d_in, A, full, full + sz, full + 2 * sz, m, n, o, ctx->stream);
    });
}

int deeds_ctx_apply_float(deeds_ctx* ctx, float* d_out, const float* d_in, const float* X, const float* d_flow,
                          int grid_step, int m, int n, int o) {
    if (!ctx) return DEEDS_ERR_ARG;
    return guarded(ctx, [&] {
        DEEDS_REQUIRE(d_out && d_in && d_flow, "null pointer");
        DEEDS_REQUIRE(d_out != d_in, "out must not alias in");
        DEEDS_REQUIRE(m > 0 && n > 0 && o > 0 && (long long)m * n * o < (1LL << 31), "bad volume dimensions");
        DEEDS_REQUIRE(grid_step >= 1 && m / grid_step >= 1 && n / grid_step >= 1 && o / grid_step >= 1,
                      "bad grid step");
        const int m1 = m / grid_step, n1 = n / grid_step, o1 = o / grid_step;
        const size_t sz1 = (size_t)m1 * n1 * o1;
        const float identity[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
        const Affine A = affine_from_X(X ? X : identity);
        // applyBCVfloat.cpp:186-198: the field upsampled in registers and the trilinear warp, one pass
        launch_warp_affine_grid(d_out, d_in, A, d_flow, d_flow + sz1, d_flow + 2 * sz1, m, n, o, m1, n1, o1,
                                ctx->stream);
    });
}

int deeds_apply_seg(int16_t* out, const int16_t* in, const float* X, const float* flow, int grid_step, int m, int n,
                    int o) {
    return guarded(nullptr, [&] {
        DEEDS_REQUIRE(out && in && flow, "null pointer");
        DEEDS_REQUIRE(grid_step >= 1 && m / grid_step >= 1 && n / grid_step >= 1 && o / grid_step >= 1,
                      "bad grid step");
        deeds_ctx* c = default_ctx();
        DEEDS_HIP(hipSetDevice(c->device));
        const size_t sz = (size_t)m * n * o;
        const size_t sz1 = (size_t)(m / grid_step) * (n / grid_step) * (o / grid_step);
        float* f = c->get<float>("h.apply.f", 3 * sz1);
        int16_t* s = c->get<int16_t>("h.apply.s", 2 * sz);
        h2d(f, flow, 3 * sz1, c->stream);
        h2d(s, in, sz, c->stream);
        int rc = deeds_ctx_apply_seg(c, s + sz, s, X, f, grid_step, m, n, o);
        if (rc) fail(rc, c->err);
        d2h(out, s + sz, sz, c->stream);
        DEEDS_HIP(hipStreamSynchronize(c->stream));
    });
}

int deeds_apply_float(float* out, const float* in, const float* X, const float* flow, int grid_step, int m, int n,
                      int o) {
    return guarded(nullptr, [&] {
        DEEDS_REQUIRE(out && in && flow, "null pointer");
        DEEDS_REQUIRE(grid_step >= 1 && m / grid_step >= 1 && n / grid_step >= 1 && o / grid_step >= 1,
                      "bad grid step");
        deeds_ctx* c = default_ctx();
        DEEDS_HIP(hipSetDevice(c->device));
        const size_t sz = (size_t)m * n * o;
        const size_t sz1 = (size_t)(m / grid_step) * (n / grid_step) * (o / grid_step);
        float* f = c->get<float>("h.apply.f", 3 * sz1);
        float* im = c->get<float>("h.apply.im", 2 * sz);
        h2d(f, flow, 3 * sz1, c->stream);
        h2d(im, in, sz, c->stream);
        int rc = deeds_ctx_apply_float(c, im + sz, im, X, f, grid_step, m, n, o);
        if (rc) fail(rc, c->err);
        d2h(out, im + sz, sz, c->stream);
        DEEDS_HIP(hipStreamSynchronize(c->stream));
    });
}

}  // extern "C"
