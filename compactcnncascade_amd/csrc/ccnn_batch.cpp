// Batch path of the MI355X detection path: every pyramid level of every frame of a batch
// (frames of any sizes and channel counts) runs as ONE graph launch:
//
//   [k_resize_u8_tab per pre-resized frame]  gray_image_only 8-bit pre-resize
//   k_gray_batch                             gray of every frame
//   k_stage1_l1, k_stage1                    stage 1 over the (frame, level, tile) work list
//   k_cmp_count, k_cmp_scan, k_cmp_emit      ordered compaction of the windows above t1
//   k_check                                  stages 2/3, a persistent grid that takes
//                                            (candidate, crop variant) items from a queue
//
// The graph is captured once per batch layout (frame sizes, channels, plans) and replayed:
// per call only the frame table (the frames' device addresses) and, for host frames, the
// uploads are issued before the launch. The candidates come out in the reference's visit
// order (frame, level in check order, raster), so the host replays each frame's
// order-dependent decisions and Merger directly (cnn_detector_v3.cpp:1842-2011, 2745-2936),
// as the per-frame path does.
#include "ccnn_detector.h"

#include <algorithm>
#include <atomic>
#include <climits>
#include <cstdio>
#include <cstring>
#include <thread>
#include <sched.h>

namespace ccnn {

#define HIP_OK(expr)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) {                                                              \
            note_error(e_);                                                                  \
            std::fprintf(stderr, "[ccnn] HIP error %s at %s:%d\n", hipGetErrorString(e_),     \
                         __FILE__, __LINE__);                                                \
            return -1;                                                                       \
        }                                                                                    \
    } while (0)

namespace {

size_t round_up(size_t v, size_t m) { return (v + m - 1) / m * m; }

}  // namespace

ReplayPool::ReplayPool(int threads)
{
    for (int t = 0; t < threads; ++t) threads_.emplace_back([this] { loop(); });
}

ReplayPool::~ReplayPool()
{
    {
        std::lock_guard<std::mutex> g(m_);
        stop_ = true;
    }
    wake_.notify_all();
    for (auto& t : threads_) t.join();
}

void ReplayPool::loop()
{
    long long seen = 0;
    for (;;) {
        const std::function<void(int)>* fn;
        int n;
        {
            std::unique_lock<std::mutex> l(m_);
            wake_.wait(l, [&] { return stop_ || gen_ != seen; });
            if (stop_) return;
            seen = gen_;
            // a batch that ended before this thread woke has no job left (fn_ cleared);
            // otherwise the thread counts as busy before run() can see the batch done
            fn = fn_;
            n = n_;
            if (!fn) continue;
            busy_++;
        }
        for (int i; (i = next_.fetch_add(1)) < n;) (*fn)(i);
        {
            std::lock_guard<std::mutex> g(m_);
            if (--busy_ == 0) done_.notify_all();
        }
    }
}

void ReplayPool::run(int n, const std::function<void(int)>& fn, const std::function<void()>& mine)
{
    {
        std::lock_guard<std::mutex> g(m_);
        fn_ = &fn;
        n_ = n;
        next_.store(0);
        gen_++;
    }
    wake_.notify_all();
    mine();
    for (int i; (i = next_.fetch_add(1)) < n;) fn(i);
    // every item is taken; wait for the threads still running one, then retire the job
    std::unique_lock<std::mutex> l(m_);
    done_.wait(l, [&] { return busy_ == 0; });
    fn_ = nullptr;
}

void Detector::release_batch()
{
    BatchArena& B = batch_;
    if (B.st) (void)hipStreamSynchronize(B.st);
    if (B.exec) { (void)hipGraphExecDestroy(B.exec); B.exec = nullptr; }
    if (B.graph) { (void)hipGraphDestroy(B.graph); B.graph = nullptr; }
    dfree(B.d_frames); dfree(B.d_pre); dfree(B.d_gray); dfree(B.d_s1); dfree(B.d_resp); dfree(B.d_levels);
    dfree(B.d_l1_order); dfree(B.d_ftab); dfree(B.d_counts); dfree(B.d_offs); dfree(B.d_meta); dfree(B.d_cands);
    dfree(B.d_s2); dfree(B.d_s3); dfree(B.d_snf); dfree(B.d_sn); dfree(B.d_keep);
    hfree(B.h_ftab); hfree(B.h_meta); hfree(B.h_cands); hfree(B.h_s23);
    for (auto& ev : B.ev)
        if (ev) (void)hipEventDestroy(ev);
    if (B.st) { (void)hipStreamDestroy(B.st); B.st = nullptr; }
    const int gen = B.gen + 1;
    B = BatchArena();
    B.gen = gen;
}

int Detector::detect_batch_graph(int n, const ccnn_image* imgs, bool device_ptr,
                                 std::vector<std::vector<ccnn_face>>& faces)
{
    faces.assign(std::max(n, 0), std::vector<ccnn_face>());
    std::vector<BatchFrame> seg;
    auto flush = [&]() -> int {
        if (seg.empty()) return 0;
        int r = run_batch(seg, device_ptr, faces);
        if (r == 1) {
            // more candidates than the batch buffers hold, or a batch too large to index:
            // the frames go one at a time through the per-frame path with their plans
            r = 0;
            Slot& s = slots_[0];
            for (const BatchFrame& f : seg) {
                ccnn_image im{f.w, f.h, f.ch, const_cast<unsigned char*>(f.data), f.step};
                if (s.busy) (void)hipStreamSynchronize(s.st);
                if (enqueue_plan(s, im, device_ptr, f.plan, f.pre, f.gw, f.gh) < 0 || finish(s, faces[f.idx]) < 0) {
                    r = -1;
                    break;
                }
            }
        }
        for (const BatchFrame& f : seg) const_cast<Plan*>(f.plan)->pins--;
        seg.clear();
        return r;
    };
    for (int i = 0; i < n; ++i) {
        const ccnn_image& im = imgs[i];
        // the guards may re-initialise the detector, which drops the plans: run the frames
        // planned so far first
        const bool guard = empty_ || !built_ || im.cols > param.max_image_width || im.rows > param.max_image_height ||
                           (gray_only_ && pre_scale_ > 0.f && im.channels != 1);
        if (guard && flush() < 0) return -1;
        const int r = prepare(im);
        if (r <= 0) continue;                   // no faces (r == 0) or a rejected frame (r < 0)
        BatchFrame f;
        f.idx = i;
        f.data = im.data;
        f.w = im.cols;
        f.h = im.rows;
        f.ch = im.channels;
        f.step = im.step;
        f.plan = plan_frame(im, &f.pre, &f.gw, &f.gh, shape_);
        const_cast<Plan*>(f.plan)->pins++;
        seg.push_back(f);
    }
    return flush();
}

// CPUs this process may run on (its affinity mask), for the replay pool's size
static int host_cpus()
{
    cpu_set_t set;
    CPU_ZERO(&set);
    if (sched_getaffinity(0, sizeof(set), &set) == 0) return CPU_COUNT(&set);
    return (int)std::thread::hardware_concurrency();
}

int Detector::run_batch(const std::vector<BatchFrame>& seg, bool device_ptr, std::vector<std::vector<ccnn_face>>& faces)
{
    HIP_OK(hipSetDevice(device_));
    BatchArena& B = batch_;
    const int nf = (int)seg.size();
    const bool checks = set_.detect_mode != CCNN_DETECT_DISABLE;

    // layout of the batch: each frame's gray image, level table, L1-sum planes, response
    // maps and stage-1 tiles follow the previous frame's
    // each frame's first level, tile, L1 tile and L1-sum offset in the batch (the level
    // table itself, the frames' levels with these offsets, is built when the layout changes)
    std::vector<int> tile_base(nf), l1_base(nf);
    std::vector<size_t> s1_base(nf);
    int nlv = 0;
    std::vector<int> lv_begin(nf + 1, 0), gpitch(nf);
    std::vector<size_t> goff(nf), foff(nf), poff(nf), roff(nf);
    std::vector<long long> key;
    key.push_back(device_ptr);
    key.push_back(checks);
    key.push_back(check_per_cu_);
    key.push_back(precision_);
    // tile skipping applies to the parameters the network was trained for
    const bool skip = d_skip_ && param.min_face_height == skip_min_face_ && param.scale_factor == skip_sf_;
    key.push_back(skip);
    for (int k = 0; k < CCNN_SKIP_SLOTS; ++k) key.push_back(skip ? (long long)(skip_threshold_[k] * 1e6f) : 0);
    key.push_back(skip ? skip_srad_ * 64 + skip_margin_ : 0);
    for (const BatchFrame& f : seg)
        if (f.plan->shape != seg[0].plan->shape) return 1;   // one tile shape per launch
    size_t gray_n = 0, s1_n = 0, resp_n = 0, frames_n = 0, pre_n = 0;
    long long tiles = 0, l1tiles = 0, quads = 0, cells = 0;
    std::vector<SnFrame> snf(nf);
    for (int fi = 0; fi < nf; ++fi) {
        const BatchFrame& f = seg[fi];
        const Plan& P = *f.plan;
        const int gp = (f.gw + 63) / 64 * 64;
        gpitch[fi] = gp;
        goff[fi] = gray_n;
        gray_n += (size_t)gp * (f.gh + 1);                 // a multiple of 64 floats
        if (!device_ptr) {
            foff[fi] = frames_n;
            frames_n += round_up(f.step * (size_t)(f.h - 1) + (size_t)f.w * f.ch, 256);
        }
        if (f.pre) {
            poff[fi] = pre_n;
            pre_n += round_up((size_t)f.gw * f.gh, 256);
        }
        lv_begin[fi] = nlv;
        nlv += (int)P.levels.size();
        tile_base[fi] = (int)tiles;
        l1_base[fi] = (int)l1tiles;
        s1_base[fi] = s1_n;
        roff[fi] = resp_n;
        resp_n += P.resp_total;
        s1_n += P.s1_total;
        tiles += P.ntiles;
        l1tiles += P.nl1tiles;
        quads += (long long)((f.gw + 15) / 16) * f.gh;
        snf[fi].h8 = (f.gh + 7) / 8;
        snf[fi].w8 = (f.gw + 7) / 8;
        snf[fi].off = snf[fi].pos_off = (int)cells;
        cells += (long long)snf[fi].h8 * snf[fi].w8;
        key.push_back(P.id);
        key.push_back(f.w);
        key.push_back(f.h);
        key.push_back(f.ch);
        key.push_back((long long)f.step);
        key.push_back(f.pre);
        // a frame whose offsets do not fit the kernels' 32-bit indices goes per frame
        if (resp_n > (size_t)INT_MAX / 2 || s1_n > (size_t)INT_MAX || tiles > INT_MAX || l1tiles > INT_MAX ||
            quads > INT_MAX)
            return 1;
    }
    lv_begin[nf] = nlv;
    const int nl = nlv;
    if (nl == 0) {
        // no level to run in any frame: no candidates
        for (const BatchFrame& f : seg) faces[f.idx].clear();
        return 0;
    }
    const int nb = (int)((resp_n + CMP_BLOCK - 1) / CMP_BLOCK);
    // candidate capacity: the compaction stores this many, the check kernel evaluates them;
    // a batch with more takes the per-frame path
    const size_t cap = std::min<size_t>(std::max<size_t>(resp_n / 64, 65536), (size_t)1 << 22);

    bool moved = false;
    auto grow = [&](auto*& p, size_t& have, size_t need) -> int {
        if (need <= have) return 0;
        if (dalloc(p, need) < 0) { have = 0; return -1; }
        have = need;
        moved = true;
        return 0;
    };
    size_t dummy_cap = 0;
    if ((!device_ptr && grow(B.d_frames, B.frames_cap, frames_n) < 0) || grow(B.d_pre, B.pre_cap, pre_n) < 0 ||
        grow(B.d_gray, B.gray_cap, gray_n) < 0 || grow(B.d_s1, B.s1_cap, s1_n) < 0 ||
        grow(B.d_resp, B.resp_cap, resp_n) < 0 || grow(B.d_levels, B.levels_cap, (size_t)nl) < 0 ||
        grow(B.d_l1_order, B.l1_cap, (size_t)(l1tiles + tiles)) < 0 ||
        grow(B.d_counts, B.cnt_cap, (size_t)nb + 1) < 0)
        return -1;
    if (moved) {
        // the offsets buffer follows the counts buffer's capacity
        if (dalloc(B.d_offs, B.cnt_cap) < 0) return -1;
    }
    if (!B.d_meta) {
        if (dalloc(B.d_meta, 4) < 0 || halloc(B.h_meta, B.hd_meta, 4) < 0) return -1;
        moved = true;
    }
    if (cap > B.cand_cap) {
        if (dalloc(B.d_cands, cap) < 0 || dalloc(B.d_s2, 2 * cap) < 0 || dalloc(B.d_s3, 2 * cap) < 0 ||
            halloc(B.h_cands, B.hd_cands, cap) < 0 || halloc(B.h_s23, B.hd_s23, 4 * cap) < 0)
            return -1;
        B.cand_cap = B.h_cap = cap;
        moved = true;
    }
    if ((size_t)nf > B.ftab_cap) {
        if (dalloc(B.d_ftab, (size_t)nf) < 0 || halloc(B.h_ftab, B.hd_ftab, (size_t)nf) < 0 ||
            dalloc(B.d_snf, (size_t)nf) < 0)
            return -1;
        B.ftab_cap = nf;
        moved = true;
    }
    if (skip) {
        if (grow(B.d_sn, B.sn_cap, (size_t)cells * (1 + 8 + 8 + CCNN_SKIP_SLOTS)) < 0 ||
            grow(B.d_keep, B.keep_cap, (size_t)(tiles + l1tiles)) < 0)
            return -1;
    }
    (void)dummy_cap;
    if (moved) {
        B.gen++;
        B.key.clear();
    }
    key.push_back(B.gen);
    key.push_back((long long)B.cand_cap);

    // frame table: the frames' addresses change from call to call, so it is uploaded
    // before every launch (the graph reads it from fixed device memory)
    long long q0 = 0;
    for (int fi = 0; fi < nf; ++fi) {
        const BatchFrame& f = seg[fi];
        FrameDesc F{};
        const uint8_t* src = device_ptr ? f.data : B.d_frames + foff[fi];
        if (f.pre) {
            F.orig = src;
            F.ow = f.w;
            F.oh = f.h;
            F.ostep = (int)f.step;
            F.src = B.d_pre + poff[fi];
            F.w = f.gw;
            F.h = f.gh;
            F.ch = 1;
            F.step = f.gw;
        } else {
            F.src = src;
            F.w = f.w;
            F.h = f.h;
            F.ch = f.ch;
            F.step = (int)f.step;
        }
        F.gray_off = (long long)goff[fi];
        F.gp = gpitch[fi];
        F.qw = (F.w + 15) / 16;
        F.quads_off = (int)q0;
        q0 += (long long)F.qw * F.h;
        const uintptr_t al = F.ch == 4 ? 16 : 4;
        F.aligned = (((uintptr_t)F.src % al) == 0 && (uintptr_t)F.step % al == 0 ? 1 : 0) |
                    (((uintptr_t)F.src % 16) == 0 && (uintptr_t)F.step % 16 == 0 ? 2 : 0);
        B.h_ftab[fi] = F;
    }
    HIP_OK(hipMemcpyAsync(B.d_ftab, B.h_ftab, sizeof(FrameDesc) * nf, hipMemcpyHostToDevice, B.st));
    if (!device_ptr)
        for (int fi = 0; fi < nf; ++fi) {
            const BatchFrame& f = seg[fi];
            const size_t bytes = f.step * (size_t)(f.h - 1) + (size_t)f.w * f.ch;
            HIP_OK(hipMemcpyAsync(B.d_frames + foff[fi], f.data, bytes, hipMemcpyHostToDevice, B.st));
        }

    // the launch sequence of the batch; ev (optional): events recorded after each group of
    // kernels (gray, proposal network, L1 sums, stage-1 tiles, compaction, checks), ev[0]
    // before the first
    auto launch_all = [&](hipEvent_t* ev, int* nlaunch) -> hipError_t {
        int launches = 0;
        hipError_t e = hipSuccess;
        auto mark = [&](int k) {
            if (ev && e == hipSuccess) e = hipEventRecord(ev[k], B.st);
        };
        mark(0);
        for (int fi = 0; fi < nf && e == hipSuccess; ++fi)
            if (seg[fi].pre) {
                e = launch_resize_u8_tab(B.d_ftab, fi, seg[fi].gw, seg[fi].gh, B.st);
                launches++;
            }
        if (e == hipSuccess) {
            e = launch_gray_batch(B.d_ftab, nf, (int)quads, B.d_gray, B.st);
            launches++;
        }
        mark(1);
        unsigned char* keep = nullptr;
        unsigned char* l1_keep = nullptr;
        if (e == hipSuccess && skip && tiles > 0) {
            // the proposal network on every frame, then the keep flags of the stage-1 tiles
            keep = B.d_keep;
            l1_keep = B.d_keep + tiles;
            float* x0 = B.d_sn;
            float* a1 = x0 + cells;
            float* a2 = a1 + 8 * cells;
            float* z = a2 + 8 * cells;
            e = launch_skipnet(B.d_gray, B.d_ftab, B.d_snf, nf, (int)cells, d_skip_, x0, a1, a2, z, B.d_levels, nl,
                               (int)tiles, (int)l1tiles, seg[0].plan->shape, skip_threshold_, skip_srad_,
                               skip_margin_, keep, l1_keep, B.st);
            launches += 6;
        }
        mark(2);
        if (e == hipSuccess && tiles > 0) {
            // with events, the two stage-1 kernels are timed apart (kev: start/stop pairs)
            e = launch_stage1(B.d_gray, 0, 0, 0, B.d_levels, nl, (int)l1tiles, (int)tiles, d_w_[0], B.d_s1, B.d_resp,
                              set_.t1, nullptr, nullptr, nullptr, 0, B.d_l1_order, B.st, ev ? ev + 7 : nullptr,
                              seg[0].plan->shape, half_stage1(), keep, l1_keep);
            launches += 2;
        } else if (ev) {
            for (int k = 7; k < 11 && e == hipSuccess; ++k) e = hipEventRecord(ev[k], B.st);
        }
        mark(3);
        if (e == hipSuccess) {
            e = launch_compact(B.d_resp, (int)resp_n, set_.t1, B.d_counts, B.d_offs, B.d_levels, nl, B.d_cands,
                               (int)B.cand_cap, B.hd_cands, (int)B.h_cap, B.d_meta, B.hd_meta, B.st);
            launches += nb > 0 ? 3 : 1;
        }
        mark(4);
        if (e == hipSuccess && checks) {
            // a persistent grid of check_per_cu workgroups per CU takes the items from the
            // queue (meta[1], zeroed by the compaction)
            e = launch_check(B.d_gray, 0, 0, 0, B.d_levels, B.d_cands, B.d_meta, check_params(), d_w_[1], d_w_[2],
                             B.d_s2, B.d_s3, nullptr, check_per_cu_ * cus_, nullptr, nullptr, nullptr, B.hd_s23,
                             B.hd_s23 + 2 * B.h_cap, (int)B.h_cap, B.st, nullptr, (int)B.cand_cap, B.d_meta + 1,
                             check_precision(), d_mw_[0], d_mw_[1]);
            launches++;
        }
        mark(5);
        if (nlaunch) *nlaunch = launches;
        return e;
    };
    const bool fresh = key != B.key;
    if (fresh) {
        // the level table and the tile tables of this layout (built only here: a replayed
        // layout keeps them on the device): every frame's L1 tiles {level, tile} in
        // XCD-band order, then every frame's stage-1 tiles in launch order
        std::vector<int2> order;
        order.reserve((size_t)(l1tiles + tiles));
        for (int pass = 0; pass < 2; ++pass)
            for (int fi = 0; fi < nf; ++fi) {
                const Plan& P = *seg[fi].plan;
                const size_t k0 = pass == 0 ? 0 : (size_t)P.nl1tiles;
                const size_t k1 = pass == 0 ? (size_t)P.nl1tiles : P.l1_order.size();
                for (size_t k = k0; k < k1; ++k)
                    order.push_back(make_int2(P.l1_order[k].x + lv_begin[fi], P.l1_order[k].y));
            }
        std::vector<LevelDesc> lv;
        lv.reserve((size_t)nl);
        for (int fi = 0; fi < nf; ++fi) {
            const BatchFrame& f = seg[fi];
            for (LevelDesc L : f.plan->levels) {
                L.frame = fi;
                L.gp = gpitch[fi];
                L.gw = f.gw;
                L.gh = f.gh;
                L.gray_off = (long long)goff[fi];
                L.resp_off += (int)roff[fi];
                L.tile_off += tile_base[fi];
                L.l1_tile_off += l1_base[fi];
                L.s1_off += (int)s1_base[fi];
                lv.push_back(L);
            }
        }
        HIP_OK(hipMemcpyAsync(B.d_levels, lv.data(), sizeof(LevelDesc) * nl, hipMemcpyHostToDevice, B.st));
        if (!order.empty())
            HIP_OK(hipMemcpyAsync(B.d_l1_order, order.data(), sizeof(int2) * order.size(), hipMemcpyHostToDevice,
                                  B.st));
        HIP_OK(hipMemcpyAsync(B.d_snf, snf.data(), sizeof(SnFrame) * nf, hipMemcpyHostToDevice, B.st));
        HIP_OK(hipStreamSynchronize(B.st));
        if (B.exec) { (void)hipGraphExecDestroy(B.exec); B.exec = nullptr; }
        if (B.graph) { (void)hipGraphDestroy(B.graph); B.graph = nullptr; }
        B.key.clear();
    }
    if (profiling_) {
        // profiled batches run the same kernels without the graph, with events between the
        // groups, so that each group's time is its own (the tables above are uploaded for
        // this layout; the graph, if any, stays as it is)
        if (!B.ev[0])
            for (auto& ev : B.ev) HIP_OK(hipEventCreate(&ev));
        int launches = 0;
        HIP_OK(launch_all(B.ev, &launches));
        HIP_OK(hipStreamSynchronize(B.st));
        const int pairs[6][2] = {{0, 1}, {1, 2}, {7, 8}, {9, 10}, {3, 4}, {4, 5}};
        for (int k = 0; k < 6; ++k) {
            float t = 0.f;
            HIP_OK(hipEventElapsedTime(&t, B.ev[pairs[k][0]], B.ev[pairs[k][1]]));
            B.ms[k] = t;
        }
        float t = 0.f;
        HIP_OK(hipEventElapsedTime(&t, B.ev[0], B.ev[5]));
        B.ms[6] = t;
        B.ms[7] = (float)nf;
        B.launches = launches;
    } else {
        if (fresh || !B.exec) {
            HIP_OK(hipStreamBeginCapture(B.st, hipStreamCaptureModeThreadLocal));
            int launches = 0;
            const hipError_t e = launch_all(nullptr, &launches);
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(B.st, &g);
            if (e != hipSuccess || ec != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                HIP_OK(e != hipSuccess ? e : ec);
            }
            B.graph = g;
            HIP_OK(hipGraphInstantiate(&B.exec, g, nullptr, nullptr, 0));
            B.key = key;
            B.launches = launches;
            batch_captures_++;
        }
        HIP_OK(hipGraphLaunch(B.exec, B.st));
        HIP_OK(hipStreamSynchronize(B.st));
    }

    B.tiles = tiles;
    B.skipped = skip && tiles > 0;
    const int total = B.h_meta[0];
    if (total < 0 || (size_t)total > B.cand_cap) return 1;
    // each frame's candidates (in check order) through the decisions, the replay and Merger;
    // the frames are independent, so host threads take them in turn
    const Candidate* C = B.h_cands;
    const int2* S2 = B.h_s23;
    const int2* S3 = B.h_s23 + 2 * B.h_cap;
    const CheckParams P = check_params();
    std::vector<int> cbeg(nf + 1, 0);
    for (int fi = 0, c = 0; fi < nf; ++fi) {
        cbeg[fi] = c;
        while (c < total && C[c].level < lv_begin[fi + 1]) ++c;
        cbeg[fi + 1] = c;
    }
    std::vector<Candidate> last_c;
    std::vector<CheckResult> last_r;
    std::vector<int> last_s;
    auto run_frame = [&](int fi, std::vector<Candidate>& lc, std::vector<CheckResult>& lr, std::vector<int>* st) {
        const BatchFrame& f = seg[fi];
        const int l0 = lv_begin[fi], l1 = lv_begin[fi + 1];
        const int c0 = cbeg[fi], cnt = cbeg[fi + 1] - c0;
        lc.resize(cnt);
        lr.resize(checks ? cnt : 0);
        for (int k = 0; k < cnt; ++k) {
            lc[k] = C[c0 + k];
            lc[k].level -= l0;
            if (checks)
                lr[k] = check_decision(P, f.plan->levels[lc[k].level].scale, S2 + 2 * (size_t)(c0 + k),
                                       S3 + 2 * (size_t)(c0 + k));
        }
        // the replay reads the levels' scales, which the frame's plan holds
        const std::vector<Det> rects = replay(f.plan->levels.data(), l1 - l0, lc.data(), checks ? lr.data() : nullptr,
                                              cnt, checks, st);
        emit_faces(rects, f.pre, faces[f.idx]);
    };
    if (!pool_) pool_.reset(new ReplayPool(std::max(0, std::min(host_cpus(), 16) - 1)));
    const std::function<void(int)> one = [&](int fi) {
        thread_local std::vector<Candidate> lc;
        thread_local std::vector<CheckResult> lr;
        run_frame(fi, lc, lr, nullptr);
    };
    // the last frame on this thread, with its record; the others over the pool
    pool_->run(nf - 1, one, [&] { run_frame(nf - 1, last_c, last_r, &last_s); });
    const BatchFrame& fl = seg[nf - 1];
    last_levels_ = fl.plan->levels;
    last_cands_.swap(last_c);
    last_results_.swap(last_r);
    last_status_.swap(last_s);
    last_count_ = (int)last_cands_.size();
    last_gray_ = B.d_gray + goff[nf - 1];
    last_gp_ = gpitch[nf - 1];
    last_gw_ = fl.gw;
    last_gh_ = fl.gh;
    last_resp_ = B.d_resp + roff[nf - 1];
    return 0;
}

void Detector::skip_stats(long long* tiles, long long* kept)
{
    *tiles = *kept = batch_.tiles;
    if (!batch_.skipped || !batch_.d_keep) return;
    std::vector<unsigned char> h((size_t)batch_.tiles);
    if (hipSetDevice(device_) != hipSuccess ||
        hipMemcpy(h.data(), batch_.d_keep, h.size(), hipMemcpyDeviceToHost) != hipSuccess)
        return;
    long long k = 0;
    for (unsigned char c : h) k += c;
    *kept = k;
}

int Detector::set_skipnet(const float* params, float threshold, int min_face, float scale_factor)
{
    dfree(d_skip_);
    batch_.key.clear();
    if (!params) return 0;
    if (!(threshold > 0.f && threshold < 1.f) || min_face <= 0 || !(scale_factor > 1.f)) return -1;
    // the inference weights: the trained ones rounded to FP16
    std::vector<float> w(params, params + CCNN_SKIP_PARAMS);
    for (float& v : w) v = (float)(_Float16)v;
    if (dalloc(d_skip_, w.size()) < 0) return -1;
    HIP_OK(hipMemcpy(d_skip_, w.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice));
    for (float& t : skip_threshold_) t = threshold;
    skip_min_face_ = min_face;
    skip_sf_ = scale_factor;
    return 0;
}

int Detector::autotune(int n, const ccnn_image* imgs, bool device_ptr, int reps, float* ms, float* grid_ms)
{
    if (n <= 0 || reps <= 0) return -1;
    const int keep_shape = shape_, keep_grid = check_per_cu_, keep_mode = batch_mode_;
    batch_mode_ = 1;
    std::vector<std::vector<ccnn_face>> faces;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return -1;
    // one batch pass, timed on the batch stream with events (the graph launch, its uploads
    // and the host replay are all inside); the detector state (num_scales) is the same for
    // every pass, so each shape sees the same work
    auto time_pass = [&](float* best) -> int {
        *best = -1.f;
        if (detect_batch(n, imgs, device_ptr, faces) < 0) return -1;      // capture + warm-up
        for (int r = 0; r < reps; ++r) {
            (void)hipEventRecord(e0, batch_.st);
            if (detect_batch(n, imgs, device_ptr, faces) < 0) return -1;
            (void)hipEventRecord(e1, batch_.st);
            (void)hipEventSynchronize(e1);
            float t = 0.f;
            (void)hipEventElapsedTime(&t, e0, e1);
            if (*best < 0.f || t < *best) *best = t;
        }
        return 0;
    };
    int best_shape = keep_shape;
    float best_t = -1.f;
    for (int id = 0; id < N_TILE_SHAPES; ++id) {
        shape_ = id;
        float t;
        if (time_pass(&t) < 0) t = -1.f;
        if (ms) ms[id] = t < 0.f ? __builtin_nanf("") : t;
        if (t >= 0.f && (best_t < 0.f || t < best_t)) { best_t = t; best_shape = id; }
    }
    shape_ = best_shape;
    int best_grid = keep_grid;
    best_t = -1.f;
    for (int g = 2; g <= 4; ++g) {
        check_per_cu_ = g;
        float t;
        if (time_pass(&t) < 0) t = -1.f;
        if (grid_ms) grid_ms[g - 2] = t < 0.f ? __builtin_nanf("") : t;
        if (t >= 0.f && (best_t < 0.f || t < best_t)) { best_t = t; best_grid = g; }
    }
    check_per_cu_ = best_grid;
    batch_mode_ = keep_mode;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best_shape;
}

}  // namespace ccnn
