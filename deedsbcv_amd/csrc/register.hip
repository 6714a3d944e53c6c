// The symmetric multi-level registration driver (reference: deedsBCV0.cpp:139-356).
//
// Everything stays resident in HBM; the host only builds the MST trees.  The
// trees depend on nothing but the fixed image, the affine-warped moving image
// and the grid step (primsGraph is called on im1b and warped0 at every level,
// deedsBCV0.cpp:275 and :302), so all of them are started at the beginning:
// the GPU computes every level's edge SADs, the host threads turn them into
// trees while the GPU runs MIND and the data costs.  Within a level the
// forward and backward directions are independent until consistentMappingCL
// (:311) and run on two streams.
#include <chrono>
#include <cmath>
#include <cstring>
#include <future>
#include <memory>
#include <string>
#include <thread>

#include "common.h"
#include "kernels.h"
#include "mst_host.h"
#include "stats.h"

namespace deeds {

size_t tree_stage_bytes(int N);
BPDevTree stage_tree(const Tree& t, void* pinned, void* d_stage, hipStream_t s);
void run_bp(const Tree& t, const BPDevTree& dt, float* d_cost, int16_t* d_inds, int* d_sel, const float* u0,
            const float* v0, const float* w0, float* u1, float* v1, float* w1, int hw, float quant, hipStream_t s);

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

struct PinnedBuf {
    void* p = nullptr;
    explicit PinnedBuf(size_t bytes) { DEEDS_HIP(hipHostMalloc(&p, std::max<size_t>(bytes, 64), hipHostMallocDefault)); }
    ~PinnedBuf() {
        if (p) (void)hipHostFree(p);
    }
    template <class T>
    T* as() {
        return static_cast<T*>(p);
    }
};

struct LevelCfg {
    int step, hw, quant, len, len3, m1, n1, o1, sz1, mind_qs;
    bool new_mind;
    CostPlan plan;
};

}  // namespace

void register_pair(deeds_ctx* ctx, const float* d_fixed, const float* d_moving, const float* X, int m, int n, int o,
                   const deeds_schedule& S, float* d_flow, float* d_warped, deeds_timing* timing) {
    using clock = std::chrono::steady_clock;
    const long long sz = (long long)m * n * o;
    const int L = S.levels;
    hipStream_t s0 = ctx->stream, s1 = ctx->aux_stream;

    // ---- schedule (deedsBCV0.cpp:183-187, :222-242) ----
    std::vector<LevelCfg> lv(L);
    int max_qs = 0;
    size_t max_cost = 0, max_pad = 0, max_sz1 = 0;
    for (int l = 0; l < L; l++) {
        LevelCfg& c = lv[l];
        c.step = S.grid[l];
        c.hw = S.radius[l];
        c.quant = S.quant[l];
        c.len = 2 * c.hw + 1;
        c.len3 = c.len * c.len * c.len;
        c.m1 = m / c.step; c.n1 = n / c.step; c.o1 = o / c.step;
        c.sz1 = c.m1 * c.n1 * c.o1;
        c.mind_qs = (int)std::floor(0.5f * (float)c.quant + 1.0f);
        c.new_mind = (l == 0) || (lv[l - 1].mind_qs != c.mind_qs);
        c.plan = datacost_plan(m, n, o, c.step, c.hw, (float)c.quant, S.alpha, 1);
        max_qs = std::max(max_qs, c.mind_qs);
        max_cost = std::max(max_cost, (size_t)c.sz1 * c.len3);
        max_pad = std::max(max_pad, (size_t)c.plan.mp * c.plan.np * c.plan.op);
        max_sz1 = std::max(max_sz1, (size_t)c.sz1);
    }
    DEEDS_REQUIRE(m >= 2 * max_qs + 1 && n >= 2 * max_qs + 1 && o >= 2 * max_qs + 1,
                  "volume too small for the MIND patch radius of the schedule");
    DEEDS_REQUIRE(m <= mind_max_dim() && n <= mind_max_dim() && o <= mind_max_dim(),
                  "a dimension exceeds the box filter's LDS tile limit");

    // ---- device workspace ----
    float* im1 = ctx->get<float>("r.im1", sz);
    float* im1b = ctx->get<float>("r.im1b", sz);
    float* warped0 = ctx->get<float>("r.warped0", sz);
    uint64_t* mind_mov = ctx->get<uint64_t>("r.mind.mov", sz);
    uint64_t* mind_fix = ctx->get<uint64_t>("r.mind.fix", sz);
    float* warped[2] = {ctx->get<float>("r.warped.0", sz), ctx->get<float>("r.warped.1", sz)};
    uint64_t* mind_w[2] = {ctx->get<uint64_t>("r.mindw.0", sz), ctx->get<uint64_t>("r.mindw.1", sz)};
    float* mw[2][2];
    for (int d = 0; d < 2; d++)
        for (int b = 0; b < 2; b++)
            mw[d][b] = ctx->get<float>("r.mindtmp." + std::to_string(d) + std::to_string(b), 6 * sz);
    uint64_t* mpad[2] = {ctx->get<uint64_t>("r.mpad.0", max_pad), ctx->get<uint64_t>("r.mpad.1", max_pad)};
    float* cost[2] = {ctx->get<float>("r.cost.0", max_cost), ctx->get<float>("r.cost.1", max_cost)};
    int16_t* inds[2] = {ctx->get<int16_t>("r.inds.0", max_cost), ctx->get<int16_t>("r.inds.1", max_cost)};
    int* sel[2] = {ctx->get<int>("r.sel.0", max_sz1), ctx->get<int>("r.sel.1", max_sz1)};
    // control-grid fields: previous level (fwd u,v,w, bwd u,v,w), this level's u0 per direction, new fields
    size_t sz2_0 = (size_t)(m / S.grid[0]) * (n / S.grid[0]) * (o / S.grid[0]);
    size_t fmax = std::max(max_sz1, sz2_0);
    float* prevf = ctx->get<float>("r.prevf", 6 * fmax);
    float* newf = ctx->get<float>("r.newf", 6 * fmax);
    float* g0[2] = {ctx->get<float>("r.g0.0", 3 * fmax), ctx->get<float>("r.g0.1", 3 * fmax)};
    float* cmtmp = ctx->get<float>("r.cm", 6 * fmax);
    double* jac = ctx->get<double>("r.jac", 4 * (size_t)L);
    double* ssd = ctx->get<double>("r.ssd", 2);

    Event ev_start, ev_stop, ev_init, ev_mind, ev_bwd_done, ev_sad;
    std::vector<std::unique_ptr<Event>> ev_lvl;
    for (int i = 0; i < 6 * L; i++) ev_lvl.emplace_back(new Event());
    auto EV = [&](int l, int k) { return ev_lvl[6 * l + k]->e; };
    // events bracketing every data-cost kernel launch, on its own stream
    std::vector<std::unique_ptr<Event>> ev_dc;
    for (int i = 0; i < 4 * L; i++) ev_dc.emplace_back(new Event());
    // marker on each direction's stream right before its BP is enqueued
    std::vector<std::unique_ptr<Event>> ev_pre;
    for (int i = 0; i < 2 * L; i++) ev_pre.emplace_back(new Event());
    auto DC = [&](int l, int dir, int end) { return ev_dc[4 * l + 2 * dir + end]->e; };

    DEEDS_HIP(hipEventRecord(ev_start.e, s0));
    auto t0 = clock::now();

    // ---- inputs: +1024 HU offset (deedsBCV0.cpp:139-144), affine warp (:210) ----
    const float identity[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    Affine A = affine_from_X(X ? X : identity);
    launch_offset(im1b, d_fixed, -1024.0f, sz, s0);
    launch_offset(im1, d_moving, -1024.0f, sz, s0);
    {
        // warpAffine with a zero field == the grid-fused warp from a zero 1x1x1 grid
        float* zero3 = g0[0];
        launch_zero(zero3, 3, s0);
        launch_warp_affine_grid(warped0, im1, A, zero3, zero3 + 1, zero3 + 2, m, n, o, 1, 1, 1, s0);
    }

    // ---- image statistics of both images (exact, chunks decomposed on the side streams) ----
    Event ev_w0;
    DEEDS_HIP(hipEventRecord(ev_w0.e, s0));
    const size_t stats_bytes = (StatsScratch::bytes((long long)sz) + 255) & ~(size_t)255;
    char* d_stats = static_cast<char*>(ctx->buf("r.stats", 2 * stats_bytes));
    PinnedBuf h_stats(2 * stats_bytes);
    double stats_time[2] = {0.0, 0.0};
    hipEvent_t w0_event = ev_w0.e;
    auto stats_task = [&, w0_event](int dir) {
        auto ts = std::chrono::steady_clock::now();
        hipStream_t ss = ctx->side_stream[dir];
        DEEDS_HIP(hipSetDevice(ctx->device));
        DEEDS_HIP(hipStreamWaitEvent(ss, w0_event, 0));
        StatsScratch sc;
        sc.dev = d_stats + dir * stats_bytes;
        sc.host = h_stats.as<char>() + dir * stats_bytes;
        float mean = 0, sd = 0;
        image_stats_device(dir ? warped0 : im1b, (long long)sz, sc, ss, &mean, &sd);
        stats_time[dir] = std::chrono::duration<double>(std::chrono::steady_clock::now() - ts).count();
        return sd;
    };
    std::shared_future<float> sd_fut[2] = {std::async(std::launch::async, stats_task, 0).share(),
                                           std::async(std::launch::async, stats_task, 1).share()};

    // ---- MST inputs: edge SADs of every level on both images ----
    size_t ec_total = 0;
    std::vector<size_t> ec_off(L);
    for (int l = 0; l < L; l++) {
        ec_off[l] = ec_total;
        ec_total += 6 * (size_t)lv[l].sz1;
    }
    float* d_ec = ctx->get<float>("r.ec", 2 * ec_total);
    for (int dir = 0; dir < 2; dir++)
        for (int l = 0; l < L; l++)
            launch_edge_sad(d_ec + dir * ec_total + ec_off[l], dir ? warped0 : im1b, lv[l].step, m, n, o, s0);
    PinnedBuf h_ec(2 * ec_total * sizeof(float));
    d2h(h_ec.as<float>(), d_ec, 2 * ec_total, s0);
    DEEDS_HIP(hipEventRecord(ev_sad.e, s0));
    hipEvent_t sad_event = ev_sad.e;
    std::vector<double> build_time(2 * L, 0.0);
    double wait_time = 0.0;
    std::vector<double> wait_level(L, 0.0);
    std::vector<std::future<std::unique_ptr<Tree>>> trees[2];
    for (int l = 0; l < L; l++)
        for (int dir = 0; dir < 2; dir++) {
            const float* ec = h_ec.as<float>() + dir * ec_total + ec_off[l];
            const LevelCfg c = lv[l];
            std::shared_future<float> sdf = sd_fut[dir];
            double* spent = &build_time[2 * l + dir];
            trees[dir].push_back(std::async(std::launch::async, [ec, c, sdf, spent, sad_event]() {
                // Prim's on the edge costs runs before the image statistics are
                // known; the exact weights then confirm the order (else rebuild).
                // The confirmation runs beside Prim's once the statistics arrive.
                DEEDS_HIP(hipEventSynchronize(sad_event));
                auto tb = std::chrono::steady_clock::now();
                auto check = std::async(std::launch::async, [ec, c, sdf]() {
                    return weights_distinguish_costs(ec, c.step, c.m1, c.n1, c.o1, sdf.get());
                });
                std::unique_ptr<Tree> t(new Tree());
                build_tree_shape(ec, c.step, c.m1, c.n1, c.o1, *t);
                const bool same_order = check.get();
                const float sd = sdf.get();
                if (same_order)
                    finish_tree(ec, sd, c.step, *t);
                else
                    build_tree(ec, sd, c.step, c.m1, c.n1, c.o1, *t);
                *spent = std::chrono::duration<double>(std::chrono::steady_clock::now() - tb).count();
                return t;
            }));
        }

    // ---- fields start at zero on the level-0 grid ----
    int m2 = m / S.grid[0], n2 = n / S.grid[0], o2 = o / S.grid[0];
    size_t sz2 = (size_t)m2 * n2 * o2;
    launch_zero(prevf, 6 * (long long)sz2, s0);
    DEEDS_HIP(hipEventRecord(ev_init.e, s0));
    DEEDS_HIP(hipStreamWaitEvent(s1, ev_init.e, 0));

    // pinned staging for every (level, direction) tree and a device area per direction
    std::vector<size_t> stage_off(L);
    size_t stage_total = 0, stage_max = 0;
    for (int l = 0; l < L; l++) {
        size_t b = tree_stage_bytes(lv[l].sz1);
        b = (b + 255) & ~(size_t)255;
        stage_off[l] = stage_total;
        stage_total += 2 * b;
        stage_max = std::max(stage_max, b);
    }
    PinnedBuf h_stage(stage_total);
    void* d_stage[2] = {ctx->buf("r.tree.0", stage_max), ctx->buf("r.tree.1", stage_max)};

    double hamming = 0.0;
    std::vector<int> tree_maxd(2 * L, 0);
    for (int l = 0; l < L; l++) {
        const LevelCfg& c = lv[l];
        DEEDS_HIP(hipEventRecord(EV(l, 0), s0));
        if (c.new_mind) {
            // both directions' MIND of warped0 / im1b (deedsBCV0.cpp:230-236)
            int rc = deeds_ctx_descriptor(ctx, mind_mov, warped0, m, n, o, c.mind_qs);
            if (rc) fail(rc, ctx->err);
            rc = deeds_ctx_descriptor(ctx, mind_fix, im1b, m, n, o, c.mind_qs);
            if (rc) fail(rc, ctx->err);
        }
        DEEDS_HIP(hipEventRecord(ev_mind.e, s0));
        DEEDS_HIP(hipStreamWaitEvent(s1, ev_mind.e, 0));
        // pre-BP work of both directions: upsample, warp, MIND, data cost
        for (int dir = 0; dir < 2; dir++) {
            hipStream_t s = dir ? s1 : s0;
            const float* pf = prevf + dir * 3 * sz2;
            float* u0 = g0[dir];
            size_t n1 = c.sz1;
            // upsampleDeformationsCL to this grid (:254 / :283)
            launch_upsample(u0, u0 + n1, u0 + 2 * n1, pf, pf + sz2, pf + 2 * sz2, c.m1, c.n1, c.o1, m2, n2, o2, s);
            // upsample to full resolution fused with the warp (:255-257 / :284-285)
            if (dir == 0)
                launch_warp_affine_grid(warped[0], im1, A, u0, u0 + n1, u0 + 2 * n1, m, n, o, c.m1, c.n1, c.o1, s);
            else
                launch_warp_image_grid(warped[1], im1b, u0, u0 + n1, u0 + 2 * n1, m, n, o, c.m1, c.n1, c.o1, s);
            if (dir == 0) DEEDS_HIP(hipEventRecord(EV(l, 1), s));
            launch_descriptor(mind_w[dir], warped[dir], m, n, o, c.mind_qs, mw[dir][0], mw[dir][1], s);
            if (dir == 0) DEEDS_HIP(hipEventRecord(EV(l, 2), s));
            launch_pad_codes(mpad[dir], mind_w[dir], m, n, o, c.plan.pad, s);
            DEEDS_HIP(hipEventRecord(DC(l, dir, 0), s));
            launch_datacost(cost[dir], dir ? mind_mov : mind_fix, mpad[dir], m, n, o, c.step, c.hw, c.quant, c.plan,
                            s);
            DEEDS_HIP(hipEventRecord(DC(l, dir, 1), s));
            if (dir == 0) DEEDS_HIP(hipEventRecord(EV(l, 3), s));
            hamming += (double)c.sz1 * c.len3 * (double)c.plan.samples;
        }
        // BP of both directions (regularisationCL, :276 / :303)
        for (int dir = 0; dir < 2; dir++) {
            hipStream_t s = dir ? s1 : s0;
            auto tw = clock::now();
            std::unique_ptr<Tree> t = trees[dir][l].get();
            double waited = std::chrono::duration<double>(clock::now() - tw).count();
            wait_time += waited;
            wait_level[l] += waited;
            tree_maxd[2 * l + dir] = t->maxd;
            DEEDS_HIP(hipEventRecord(ev_pre[2 * l + dir]->e, s));
            size_t half = (stage_off[l + 1 < L ? l + 1 : l] - stage_off[l]) / 2;
            if (l + 1 >= L) half = (stage_total - stage_off[l]) / 2;
            char* slot = h_stage.as<char>() + stage_off[l] + dir * half;
            BPDevTree dt = stage_tree(*t, slot, d_stage[dir], s);
            float* u0 = g0[dir];
            size_t n1 = c.sz1;
            float* nf = newf + dir * 3 * n1;
            run_bp(*t, dt, cost[dir], inds[dir], sel[dir], u0, u0 + n1, u0 + 2 * n1, nf, nf + n1, nf + 2 * n1, c.hw,
                   (float)c.quant, s);
            if (dir == 0) DEEDS_HIP(hipEventRecord(EV(l, 4), s));
        }
        DEEDS_HIP(hipEventRecord(ev_bwd_done.e, s1));
        DEEDS_HIP(hipStreamWaitEvent(s0, ev_bwd_done.e, 0));
        // consistentMappingCL (:311) on the new fields, then they become "previous"
        {
            size_t n1 = c.sz1;
            float* f[6] = {newf, newf + n1, newf + 2 * n1, newf + 3 * n1, newf + 4 * n1, newf + 5 * n1};
            launch_consistent(f, cmtmp, c.m1, c.n1, c.o1, c.step, s0);
            launch_jacobian(jac + 4 * l, newf, newf + n1, newf + 2 * n1, c.m1, c.n1, c.o1, c.step, s0);
            DEEDS_HIP(hipMemcpyAsync(prevf, newf, 6 * n1 * sizeof(float), hipMemcpyDeviceToDevice, s0));
        }
        DEEDS_HIP(hipEventRecord(EV(l, 5), s0));
        DEEDS_HIP(hipStreamWaitEvent(s1, EV(l, 5), 0));
        m2 = c.m1; n2 = c.n1; o2 = c.o1; sz2 = c.sz1;
    }
    DEEDS_HIP(hipEventRecord(ev_stop.e, s0));

    // ---- outputs (deedsBCV0.cpp:339-356) ----
    DEEDS_HIP(hipMemcpyAsync(d_flow, prevf, 3 * sz2 * sizeof(float), hipMemcpyDeviceToDevice, s0));
    float ssd_h[2] = {0, 0};
    if (d_warped) {
        float* wtmp = warped[0];
        launch_warp_affine_grid(wtmp, im1, A, prevf, prevf + sz2, prevf + 2 * sz2, m, n, o, m2, n2, o2, s0);
        launch_zero((float*)ssd, 4, s0);
        launch_ssd(ssd, im1b, wtmp, im1, sz, s0);
        launch_offset(d_warped, wtmp, 1024.0f, sz, s0);
    }
    DEEDS_HIP(hipStreamSynchronize(s0));
    DEEDS_HIP(hipStreamSynchronize(s1));
    auto t1 = clock::now();
    if (timing) {
        std::memset(timing, 0, sizeof(*timing));
        timing->levels = L;
        for (int l = 0; l < L; l++) {
            float a = 0, b = 0, c = 0, d = 0, e = 0;
            DEEDS_HIP(hipEventElapsedTime(&a, EV(l, 0), EV(l, 1)));
            DEEDS_HIP(hipEventElapsedTime(&b, EV(l, 1), EV(l, 2)));
            DEEDS_HIP(hipEventElapsedTime(&c, EV(l, 2), EV(l, 3)));
            DEEDS_HIP(hipEventElapsedTime(&d, EV(l, 3), EV(l, 4)));
            DEEDS_HIP(hipEventElapsedTime(&e, EV(l, 4), EV(l, 5)));
            timing->transf[l] = (a + e) * 1e-3;
            timing->mind[l] = b * 1e-3;
            timing->data[l] = c * 1e-3;
            timing->mst_reg[l] = d * 1e-3;
        }
        float tot = 0;
        DEEDS_HIP(hipEventElapsedTime(&tot, ev_start.e, ev_stop.e));
        timing->total = tot * 1e-3;
        timing->hamming_evals = hamming;
        timing->tree_wait_s = wait_time;
        timing->stats_s = std::max(stats_time[0], stats_time[1]);
        for (int l = 0; l < L; l++) {
            double lvl = 0.0;
            for (int dir = 0; dir < 2; dir++) {
                float ms = 0;
                DEEDS_HIP(hipEventElapsedTime(&ms, DC(l, dir, 0), DC(l, dir, 1)));
                lvl += ms * 1e-3;
            }
            timing->datacost_level_s[l] = lvl;
            timing->datacost_level_evals[l] = 2.0 * lv[l].sz1 * lv[l].len3 * (double)lv[l].plan.samples;
            timing->datacost_kernel_s += lvl;
            timing->tree_build_s += build_time[2 * l] + build_time[2 * l + 1];
            timing->tree_build_level_s[l] = std::max(build_time[2 * l], build_time[2 * l + 1]);
            timing->tree_wait_level_s[l] = wait_level[l];
            timing->datacost_launches += 2;
            for (int dir = 0; dir < 2; dir++) {
                float idle = 0;
                DEEDS_HIP(hipEventElapsedTime(&idle, DC(l, dir, 1), ev_pre[2 * l + dir]->e));
                timing->bp_idle_level_s[l] += std::max(0.0f, idle) * 1e-3;
            }
            float bp = 0;
            DEEDS_HIP(hipEventElapsedTime(&bp, ev_pre[2 * l]->e, EV(l, 4)));
            timing->bp_level_s[l] = bp * 1e-3;
        }
        std::vector<double> hj(4 * (size_t)L);
        DEEDS_HIP(hipMemcpy(hj.data(), jac, hj.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int l = 0; l < L; l++) {
            double N = (double)lv[l].sz1;
            double mean = hj[4 * l] / N;
            double var = N > 1 ? (hj[4 * l + 1] - N * mean * mean) / (N - 1) : 0.0;
            timing->jac_std[l] = (float)std::sqrt(std::max(var, 0.0));
            timing->jac_neg[l] = (float)(hj[4 * l + 2] / N);
        }
        if (d_warped) {
            double hs[2];
            DEEDS_HIP(hipMemcpy(hs, ssd, 2 * sizeof(double), hipMemcpyDeviceToHost));
            timing->ssd_before = (float)(hs[0] / (double)sz);
            timing->ssd_after = (float)(hs[1] / (double)sz);
        }
    }
    (void)t0;
    (void)t1;
    (void)ssd_h;
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

int deeds_apply_seg(int16_t* out, const int16_t* in, const float* X, const float* flow, int grid_step, int m, int n,
                    int o) {
    return guarded(nullptr, [&] {
        DEEDS_REQUIRE(out && in && flow, "null pointer");
        DEEDS_REQUIRE(grid_step >= 1 && m / grid_step >= 1 && n / grid_step >= 1 && o / grid_step >= 1,
                      "bad grid step");
        deeds_ctx* c = default_ctx();
        DEEDS_HIP(hipSetDevice(c->device));
        size_t sz = (size_t)m * n * o;
        int m1 = m / grid_step, n1 = n / grid_step, o1 = o / grid_step;
        size_t sz1 = (size_t)m1 * n1 * o1;
        float* f = c->get<float>("h.apply.f", 3 * sz1 + 3 * sz);
        int16_t* s = c->get<int16_t>("h.apply.s", 2 * sz);
        h2d(f, flow, 3 * sz1, c->stream);
        h2d(s, in, sz, c->stream);
        float* full = f + 3 * sz1;
        launch_upsample(full, full + sz, full + 2 * sz, f, f + sz1, f + 2 * sz1, m, n, o, m1, n1, o1, c->stream);
        const float identity[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
        Affine A = affine_from_X(X ? X : identity);
        launch_warp_affine_s(s + sz, s, A, full, full + sz, full + 2 * sz, m, n, o, c->stream);
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
        size_t sz = (size_t)m * n * o;
        int m1 = m / grid_step, n1 = n / grid_step, o1 = o / grid_step;
        size_t sz1 = (size_t)m1 * n1 * o1;
        float* f = c->get<float>("h.apply.f", 3 * sz1);
        float* im = c->get<float>("h.apply.im", 2 * sz);
        h2d(f, flow, 3 * sz1, c->stream);
        h2d(im, in, sz, c->stream);
        const float identity[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
        Affine A = affine_from_X(X ? X : identity);
        launch_warp_affine_grid(im + sz, im, A, f, f + sz1, f + 2 * sz1, m, n, o, m1, n1, o1, c->stream);
        d2h(out, im + sz, sz, c->stream);
        DEEDS_HIP(hipStreamSynchronize(c->stream));
    });
}

}  // extern "C"
