// gmmhip_server_*: a continuous-batching request loop over one resident UBM
// (include/gmmhip.h).
//
// Threads: callers submit into a FIFO; one dispatcher thread forms a batch
// whenever a slot is free (admission: FIFO order, while the batch's device
// layout fits the slot's budget and request cap), stages it, enqueues its
// device pipeline on the slot's stream and hands the slot to that slot's
// completer thread, which waits for the slot's event, scatters the results
// into the callers' buffers and frees the slot.  With several slots, batch
// k+1 is staged and launched while batch k runs.
//
// Device pipeline of one batch (requests ordered MAP, MLLR, SCORE):
//   one H2D copy of [request frame ranges | chunk tables | frames],
//   FP32 form:  loglik_kernel over all frames, stats_queue_kernel over the
//               chunks of the MAP and MLLR requests only;
//   mixed form: prep_frames (z = (x - o) * s split hi/lo), loglik_mixed,
//               stats_mixed over the MAP/MLLR chunks (the persistent
//               per-XCD work-queue form where the depth allows);
//   seg_reduce (n, F per request, MAP means of the MAP requests), seg_sum
//   (LL per request), mllr_check / the two FP64 G, k GEMMs /
//   mllr_solve (MLLR requests),
//   one D2H copy of [LL | MLLR status | n | W | means | per-frame LL].
// Chunks start at each request's first frame, the mixed form's centring o
// and scaling s come from the UBM (not from the batch), and every
// per-request sum runs in a fixed order, so results do not depend on the
// batch-mates.  In the FP16 mixed form a request with a frame value outside
// the FP16 range after centring and scaling (|z| > 239, checked at
// submission) runs in an FP32 batch instead.
#include "batch.h"
#include "ctx.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>

using namespace gmmhip_detail;
using Clock = std::chrono::steady_clock;

namespace {

constexpr size_t kAlign = 256;
inline size_t al(size_t b) { return (b + kAlign - 1) / kAlign * kAlign; }
inline uint64_t pad32(uint64_t n) { return (n + 31) / 32 * 32; }

struct Pending {
    gmmhip_request *r;
    Clock::time_point t0;
};

// Sizes of one batch.  N counts frame rows, including the padding that
// aligns each request of a mixed batch to 32 rows.
struct Shape {
    bool mixed = false;
    uint64_t N = 0;
    int U = 0, Umap = 0, Umllr = 0, nchunks = 0;
    bool frame_ll = false;
};

struct Dims {
    int M, D, E, NP;
    int KP, Mp;    // FP32 form: contraction depth, padded components
    int KPm, Mpm;  // mixed form
};

// Byte offsets in the slot arena.  IN [beg, end, lo, hi, first, X] and OUT
// [ll, status, n, W, means, LLnat] are contiguous so that each moves in one
// copy; the pinned host staging uses the same offsets.
struct Layout {
    size_t beg, end, lo, hi, first, X, in_end;
    size_t ll, status, n, W, means, LLnat, out_end, out_copy;
    size_t queue, LL2, blockLL, part, F, G, K;
    size_t zah, zal, bad, nrange, total;
};

Layout layout_of(const Shape &s, const Dims &d) {
    Layout L{};
    const int Us = s.Umap + s.Umllr;
    const int KP = s.mixed ? d.KPm : d.KP, Mp = s.mixed ? d.Mpm : d.Mp;
    const uint64_t Nz = s.mixed ? pad128(s.N) : s.N;
    size_t p = 0;
    auto take = [&p](size_t bytes) {
        const size_t at = p;
        p += al(bytes);
        return at;
    };
    L.beg = take((size_t)s.U * 8);
    L.end = take((size_t)s.U * 8);
    L.lo = take((size_t)s.nchunks * 8);
    L.hi = take((size_t)s.nchunks * 8);
    L.first = take((size_t)(Us + 1) * 4);
    L.X = take(s.N * d.D * 4);
    L.in_end = p;
    L.ll = take((size_t)s.U * 8);
    L.status = take((size_t)s.Umllr * 4);
    L.n = take((size_t)Us * d.M * 8);
    L.W = take((size_t)s.Umllr * d.D * d.E * 8);
    L.means = take((size_t)s.Umap * d.M * d.D * 4);
    L.LLnat = take(s.N * 4);
    L.out_end = p;
    L.out_copy = (s.frame_ll ? L.out_end : L.LLnat) - L.ll;
    L.queue = take(8 * 4);  // work-queue counters (FP32: one; mixed: one per XCD)
    L.LL2 = take(pad128(s.N) * 4);  // the mixed statistics kernel stages 32-frame steps
    L.blockLL = take(pad128(s.N) / 128 * 8);
    L.part = take((size_t)s.nchunks * Mp * KP * 4);
    L.F = take((size_t)Us * d.M * d.D * 8);
    L.G = take((size_t)s.Umllr * d.D * d.NP * 8);
    L.K = take((size_t)s.Umllr * d.D * d.E * 8);
    const size_t za = s.mixed ? Nz * KP * 2 : 0;
    L.zah = take(za);
    L.zal = take(za);
    L.bad = take(s.mixed ? s.N : 0);
    L.nrange = take(s.mixed ? 4 : 0);
    L.total = p;
    return L;
}

// Frame rows a request takes in a batch of the given form.
inline uint64_t rows_of(uint64_t n, bool mixed) { return mixed ? pad32(n) : n; }

int kind_rank(int k) { return k == GMMHIP_REQ_MAP ? 0 : (k == GMMHIP_REQ_MLLR ? 1 : 2); }

// FREE -> STAGING (dispatcher stages and launches) -> RUNNING (completer waits
// and scatters) -> FREE; STOP ends the completer.
enum SlotState { SLOT_FREE, SLOT_STAGING, SLOT_RUNNING, SLOT_STOP };

}  // namespace

struct gmmhip_server {
    gmmhip_t base = nullptr;  // owns the packed model
    int dev = 0;
    gmmhip_server_cfg cfg{};
    Dims d{};
    uint64_t cf = 512;
    int num_cus = 256;
    ModelOperands mo{};

    // mixed form: precision, UBM-based centring / scaling, split model
    int prec = GMMHIP_PREC_FP32;
    std::vector<float> h_zoff, h_zscale;
    float zlimit = INFINITY;
    float *d_zoff = nullptr, *d_zscale = nullptr;
    void *d_P16 = nullptr;

    struct Slot {
        int id = 0;
        hipStream_t stream = nullptr;
        hipEvent_t ev0 = nullptr, ev1 = nullptr;
        char *arena = nullptr;
        // pinned staging of exactly slot_bytes, allocated at creation: a
        // batch's IN region [0, in_end) and OUT region [ll, out_end) are
        // disjoint parts of one layout of at most slot_bytes, so IN is staged
        // at h_stage and OUT lands at h_stage + ll (h_in / h_out point there)
        char *h_stage = nullptr;
        size_t h_stage_cap = 0;
        char *h_in = nullptr, *h_out = nullptr;
        SlotState state = SLOT_FREE;
        std::vector<Pending> batch;
        std::vector<uint64_t> beg;  // first frame row of each request
        Shape shape;
        Layout lay{};
        int err = 0;
        std::thread completer;
    };
    std::vector<Slot> slots;

    std::mutex mu;
    std::condition_variable cv_dispatch, cv_slot, cv_done;
    std::deque<Pending> q_mixed, q_fp32;  // q_mixed is used by the mixed form only
    bool closing = false;
    std::thread dispatcher;
    std::string err;

    // statistics
    uint64_t submitted = 0, completed = 0, failed = 0, batches = 0, frames = 0;
    uint64_t batch_requests = 0, fp32_fallbacks = 0;
    double batch_gpu_ms = 0.0;
    std::vector<double> lat;
    Clock::time_point t_reset;
};

namespace {

void set_err(gmmhip_server *s, const std::string &e) {
    std::lock_guard<std::mutex> g(s->mu);
    s->err = e;
}

// Finish a request with a status (server mutex held by the caller).
void finish_locked(gmmhip_server *s, gmmhip_request *r, int status, Clock::time_point t0) {
    r->latency_ms = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
    __atomic_store_n(&r->status, status, __ATOMIC_RELEASE);
    if (status == GMMHIP_OK || status == GMMHIP_ERR_MLLR) {
        s->completed++;
        s->lat.push_back(r->latency_ms);
    } else {
        s->failed++;
    }
}

#define SRV_HIPCHK(expr)                                                                  \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) {                                                           \
            set_err(s, std::string(#expr " failed: ") + hipGetErrorString(e_));           \
            return e_ == hipErrorOutOfMemory ? GMMHIP_ERR_OOM : GMMHIP_ERR_HIP;           \
        }                                                                                 \
    } while (0)

// Stage and enqueue one batch on its slot (dispatcher thread).
int launch_batch(gmmhip_server *s, gmmhip_server::Slot &sl) {
    const Dims &d = s->d;
    std::stable_sort(sl.batch.begin(), sl.batch.end(), [](const Pending &a, const Pending &b) {
        return kind_rank(a.r->kind) < kind_rank(b.r->kind);
    });
    Shape sh = sl.shape;  // form set by the dispatcher
    sh.U = (int)sl.batch.size();
    sh.Umap = sh.Umllr = 0;
    sh.frame_ll = false;
    sl.beg.assign(sh.U, 0);
    std::vector<uint64_t> end(sh.U), lo, hi;
    std::vector<int> first;
    uint64_t row = 0;
    for (int u = 0; u < sh.U; u++) {
        const gmmhip_request *r = sl.batch[u].r;
        sl.beg[u] = row;
        end[u] = row + r->n_frames;
        row += rows_of(r->n_frames, sh.mixed);
        if (r->frame_ll) sh.frame_ll = true;
        if (r->kind == GMMHIP_REQ_SCORE) continue;
        if (r->kind == GMMHIP_REQ_MAP) sh.Umap++;
        else sh.Umllr++;
        first.push_back((int)lo.size());
        for (uint64_t f = sl.beg[u]; f < end[u]; f += s->cf) {
            lo.push_back(f);
            hi.push_back(std::min(end[u], f + s->cf));
        }
    }
    first.push_back((int)lo.size());
    sh.N = row;
    sh.nchunks = (int)lo.size();
    const Layout L = layout_of(sh, d);
    sl.shape = sh;
    sl.lay = L;
    if (L.out_end > sl.h_stage_cap) {
        // admission keeps every layout within slot_bytes
        set_err(s, "batch layout exceeds the slot's staging");
        return GMMHIP_ERR_STATE;
    }
    sl.h_in = sl.h_stage;
    sl.h_out = sl.h_stage + L.ll;

    // IN staging; the padding rows of a mixed batch are zero
    std::memcpy(sl.h_in + L.beg, sl.beg.data(), (size_t)sh.U * 8);
    std::memcpy(sl.h_in + L.end, end.data(), (size_t)sh.U * 8);
    if (sh.nchunks) {
        std::memcpy(sl.h_in + L.lo, lo.data(), lo.size() * 8);
        std::memcpy(sl.h_in + L.hi, hi.data(), hi.size() * 8);
    }
    std::memcpy(sl.h_in + L.first, first.data(), first.size() * 4);
    float *hx = (float *)(sl.h_in + L.X);
    for (int u = 0; u < sh.U; u++) {
        const gmmhip_request *r = sl.batch[u].r;
        std::memcpy(hx + sl.beg[u] * d.D, r->frames, r->n_frames * d.D * 4);
        const uint64_t pad = rows_of(r->n_frames, sh.mixed) - r->n_frames;
        if (pad) std::memset(hx + end[u] * d.D, 0, pad * d.D * 4);
    }

    char *A = sl.arena;
    hipStream_t st = sl.stream;
    SRV_HIPCHK(hipEventRecord(sl.ev0, st));
    SRV_HIPCHK(hipMemcpyAsync(A, sl.h_in, L.in_end, hipMemcpyHostToDevice, st));

    gmmhip_ctx *b = s->base;
    const int Us = sh.Umap + sh.Umllr;
    float *LL2 = (float *)(A + L.LL2), *LLnat = (float *)(A + L.LLnat);
    const uint64_t *clo = (const uint64_t *)(A + L.lo), *chi = (const uint64_t *)(A + L.hi);
    const float *roff = b->d_off, *rscale = nullptr;
    int rKP = d.KP, rMp = d.Mp;
    if (!sh.mixed) {
        EstepArgs a{};
        a.X = (const float *)(A + L.X);
        a.ldx = d.D;
        a.N = sh.N;
        a.D = d.D;
        a.KP = d.KP;
        a.P = b->d_P;
        a.M = d.M;
        a.Mt = b->Mt;
        a.off = b->d_off;
        a.min_ll = b->opt.min_loglike;
        a.min_gamma = b->opt.min_gamma;
        a.LL2 = LL2;
        a.LLnat = LLnat;
        a.blockLL = (double *)(A + L.blockLL);
        a.ngroups = (b->Mt + 3) / 4;
        a.partials = (float *)(A + L.part);
        a.chunk_frames = s->cf;
        SRV_HIPCHK(launch_loglik_fp32(a, st));
        if (Us > 0) {
            a.nchunks = sh.nchunks;
            a.chunk_lo = clo;
            a.chunk_hi = chi;
            a.queue = (unsigned int *)(A + L.queue);
            a.queue_blocks = std::min(sh.nchunks * a.ngroups, 2 * s->num_cus);
            SRV_HIPCHK(launch_fill_u32(a.queue, 0u, 1, st));
            SRV_HIPCHK(launch_stats_fp32_queue(a, st));
        }
    } else {
        const uint64_t Nz = pad128(sh.N);
        uint8_t *bad = (uint8_t *)(A + L.bad);
        unsigned int *nrange = (unsigned int *)(A + L.nrange);
        SRV_HIPCHK(launch_fill_u32(nrange, 0u, 1, st));
        PrepArgs p{};
        p.X = (const float *)(A + L.X);
        p.N = sh.N;
        p.Npad = Nz;
        p.D = d.D;
        p.KP = d.KPm;
        p.off = s->d_zoff;
        p.scale = s->d_zscale;
        p.za_hi_raw = (uint16_t *)(A + L.zah);
        p.za_lo_raw = (uint16_t *)(A + L.zal);
        p.bad = bad;
        p.nrange = nrange;
        p.zlimit = s->zlimit;
        SRV_HIPCHK(launch_prep_frames(p, s->prec, st));
        MixedArgs m{};
        m.N = sh.N;
        m.Npad = Nz;
        m.D = d.D;
        m.KP = d.KPm;
        m.za_hi = p.za_hi_raw;
        m.za_lo = p.za_lo_raw;
        m.bad = bad;
        m.P16 = s->d_P16;
        m.M = d.M;
        m.Mt32 = d.Mpm / 32;
        m.min_ll = b->opt.min_loglike;
        m.min_gamma = b->opt.min_gamma;
        m.LL2 = LL2;
        m.LLnat = LLnat;
        m.blockLL = (double *)(A + L.blockLL);
        SRV_HIPCHK(launch_loglik_mixed(m, s->prec, st));
        if (Us > 0) {
            m.chunk_frames = s->cf;
            m.nchunks = sh.nchunks;
            m.st_nw = 4;
            m.ngroups = (m.Mt32 + 3) / 4;
            m.partials = (float *)(A + L.part);
            m.chunk_lo = clo;
            m.chunk_hi = chi;
            m.queue = (unsigned int *)(A + L.queue);
            m.queue_blocks = mixed_stats_queue_blocks(d.KPm, m.st_nw, s->num_cus,
                                                      sh.nchunks * m.ngroups);
            SRV_HIPCHK(launch_fill_u32(m.queue, 0u, 8, st));
            SRV_HIPCHK(launch_stats_mixed_queue(m, s->prec, st));
        }
        roff = s->d_zoff;
        rscale = s->d_zscale;
        rKP = d.KPm;
        rMp = d.Mpm;
    }
    if (Us > 0) {
        SegReduceArgs r{};
        r.partials = (const float *)(A + L.part);
        r.chunk_first = (const int *)(A + L.first);
        r.off = roff;
        r.scale = rscale;
        r.U = Us;
        r.M = d.M;
        r.Mp = rMp;
        r.D = d.D;
        r.KP = rKP;
        r.n = (double *)(A + L.n);
        r.F = (double *)(A + L.F);
        r.mu = s->mo.mu;
        r.tau = s->cfg.tau;
        r.Umeans = sh.Umap;
        r.means = (float *)(A + L.means);
        SRV_HIPCHK(launch_seg_reduce(r, st));
    }
    SRV_HIPCHK(launch_seg_sum(LLnat, (const uint64_t *)(A + L.beg), (const uint64_t *)(A + L.end),
                              sh.U, (double *)(A + L.ll), st));
    if (sh.Umllr > 0) {
        const double *n = (const double *)(A + L.n) + (size_t)sh.Umap * d.M;
        const double *F = (const double *)(A + L.F) + (size_t)sh.Umap * d.M * d.D;
        int *status = (int *)(A + L.status);
        double *G = (double *)(A + L.G), *K = (double *)(A + L.K);
        SRV_HIPCHK(launch_mllr_check(n, sh.Umllr, d.M, d.D, status, st));
        MllrGkArgs g{};
        g.n = n;
        g.F = F;
        g.ivar = s->mo.ivar;
        g.U = sh.Umllr;
        g.M = d.M;
        g.D = d.D;
        g.B = s->mo.P;
        g.ldb = d.NP;
        g.Cn = d.NP;
        g.C = G;
        g.ldc = d.NP;
        g.kmode = 0;
        SRV_HIPCHK(launch_mllr_gk(g, st));
        g.B = s->mo.Xi;
        g.ldb = d.E;
        g.Cn = d.E;
        g.C = K;
        g.ldc = d.E;
        g.kmode = 1;
        SRV_HIPCHK(launch_mllr_gk(g, st));
        MllrSolveArgs m{};
        m.Gp = G;
        m.Kp = K;
        m.Ginit = s->cfg.mllr_prior ? s->mo.Ginit : nullptr;
        m.kinit = s->cfg.mllr_prior ? s->mo.kinit : nullptr;
        m.U = sh.Umllr;
        m.D = d.D;
        m.ldg = d.NP;
        m.ldk = d.E;
        m.status = status;
        m.W = (double *)(A + L.W);
        SRV_HIPCHK(launch_mllr_solve(m, st));
    }
    SRV_HIPCHK(hipMemcpyAsync(sl.h_out, A + L.ll, L.out_copy, hipMemcpyDeviceToHost, st));
    SRV_HIPCHK(hipEventRecord(sl.ev1, st));
    return GMMHIP_OK;
}

// Copy the results of a finished batch to the callers (completer thread).
void scatter(gmmhip_server *s, gmmhip_server::Slot &sl) {
    const Dims &d = s->d;
    const Layout &L = sl.lay;
    // the staging holds the device OUT region from L.ll on
    auto at = [&](size_t off) { return (const void *)(sl.h_out + (off - L.ll)); };
    const double *ll = (const double *)at(L.ll);
    const double *n = (const double *)at(L.n);
    const double *W = (const double *)at(L.W);
    const float *means = (const float *)at(L.means);
    const float *fll = (const float *)at(L.LLnat);
    const int Umap = sl.shape.Umap;
    for (int u = 0; u < sl.shape.U; u++) {
        gmmhip_request *r = sl.batch[u].r;
        r->ll_sum = ll[u];
        if (r->frame_ll) std::memcpy(r->frame_ll, fll + sl.beg[u], r->n_frames * 4);
        if (r->kind == GMMHIP_REQ_SCORE) continue;
        if (r->occ) std::memcpy(r->occ, n + (size_t)u * d.M, (size_t)d.M * 8);
        if (r->kind == GMMHIP_REQ_MAP) {
            std::memcpy(r->means, means + (size_t)u * d.M * d.D, (size_t)d.M * d.D * 4);
        } else {
            const int j = u - Umap;
            std::memcpy(r->mllr_W, W + (size_t)j * d.D * d.E, (size_t)d.D * d.E * 8);
        }
    }
}

void completer_main(gmmhip_server *s, int id) {
    hipSetDevice(s->dev);
    gmmhip_server::Slot &sl = s->slots[id];
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_slot.wait(lk, [&] { return sl.state == SLOT_RUNNING || sl.state == SLOT_STOP; });
            if (sl.state == SLOT_STOP) return;
        }
        int st = sl.err;
        float ms = 0.0f;
        if (!st) {
            const hipError_t e = hipEventSynchronize(sl.ev1);
            if (e != hipSuccess) {
                set_err(s, std::string("batch failed: ") + hipGetErrorString(e));
                st = GMMHIP_ERR_HIP;
            } else {
                hipEventElapsedTime(&ms, sl.ev0, sl.ev1);
                scatter(s, sl);
            }
        }
        const Layout &L = sl.lay;
        const int *status = (const int *)(sl.h_out + (L.status - L.ll));
        std::lock_guard<std::mutex> g(s->mu);
        for (int u = 0; u < (int)sl.batch.size(); u++) {
            gmmhip_request *r = sl.batch[u].r;
            int rs = st;
            if (!rs && r->kind == GMMHIP_REQ_MLLR && status[u - sl.shape.Umap] != 0)
                rs = GMMHIP_ERR_MLLR;
            finish_locked(s, r, rs, sl.batch[u].t0);
        }
        if (!st) {
            s->batches++;
            s->batch_requests += sl.batch.size();
            s->frames += sl.shape.N;
            s->batch_gpu_ms += ms;
        }
        sl.batch.clear();
        sl.err = 0;
        sl.state = SLOT_FREE;
        s->cv_done.notify_all();
        s->cv_dispatch.notify_all();
    }
}

void dispatcher_main(gmmhip_server *s) {
    hipSetDevice(s->dev);
    const uint64_t cap = s->cfg.slot_bytes;
    for (;;) {
        int id = -1;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            auto empty = [&] { return s->q_mixed.empty() && s->q_fp32.empty(); };
            s->cv_dispatch.wait(lk, [&] {
                if (empty()) return s->closing;
                for (auto &sl : s->slots)
                    if (sl.state == SLOT_FREE) return true;
                return false;
            });
            if (empty() && s->closing) {
                // wait for running batches, then stop the completers
                s->cv_dispatch.wait(lk, [&] {
                    for (auto &sl : s->slots)
                        if (sl.state != SLOT_FREE) return false;
                    return true;
                });
                for (auto &sl : s->slots) sl.state = SLOT_STOP;
                s->cv_slot.notify_all();
                return;
            }
            for (auto &sl : s->slots)
                if (sl.state == SLOT_FREE) {
                    id = sl.id;
                    break;
                }
            // one form per batch: the queue whose head waited longest
            const bool mixed =
                !s->q_mixed.empty() && (s->q_fp32.empty() || s->q_mixed.front().t0 <= s->q_fp32.front().t0);
            std::deque<Pending> &q = mixed ? s->q_mixed : s->q_fp32;
            // admission: FIFO while the batch layout fits the slot budget
            gmmhip_server::Slot &sl = s->slots[id];
            Shape sh;
            sh.mixed = mixed;
            while (!q.empty() && (int)sl.batch.size() < s->cfg.max_batch) {
                const Pending &p = q.front();
                Shape t = sh;
                t.U++;
                t.N += rows_of(p.r->n_frames, mixed);
                if (p.r->kind != GMMHIP_REQ_SCORE) {
                    (p.r->kind == GMMHIP_REQ_MAP ? t.Umap : t.Umllr)++;
                    t.nchunks += (int)((p.r->n_frames + s->cf - 1) / s->cf);
                }
                t.frame_ll = t.frame_ll || p.r->frame_ll;
                if (layout_of(t, s->d).total > cap) {
                    if (sl.batch.empty()) {
                        // can never fit
                        finish_locked(s, p.r, GMMHIP_ERR_OOM, p.t0);
                        s->err = "request larger than the slot budget";
                        q.pop_front();
                        s->cv_done.notify_all();
                        continue;
                    }
                    break;
                }
                sh = t;
                sl.batch.push_back(p);
                q.pop_front();
            }
            if (sl.batch.empty()) continue;
            sl.shape = sh;
            sl.state = SLOT_STAGING;  // owned by this thread until handed over
        }
        gmmhip_server::Slot &sl = s->slots[id];
        sl.err = launch_batch(s, sl);
        if (sl.err) {
            // let the stream drain before the completer fails the batch
            hipStreamSynchronize(sl.stream);
        }
        std::lock_guard<std::mutex> g(s->mu);
        sl.state = SLOT_RUNNING;
        s->cv_slot.notify_all();
    }
}

void free_server(gmmhip_server *s) {
    for (auto &sl : s->slots) {
        if (sl.arena) hipFree(sl.arena);
        if (sl.h_stage) hipHostFree(sl.h_stage);
        if (sl.ev0) hipEventDestroy(sl.ev0);
        if (sl.ev1) hipEventDestroy(sl.ev1);
        if (sl.stream) hipStreamDestroy(sl.stream);
    }
    if (s->d_zoff) hipFree(s->d_zoff);
    if (s->d_zscale) hipFree(s->d_zscale);
    if (s->d_P16) hipFree(s->d_P16);
    if (s->base) gmmhip_destroy(s->base);
    delete s;
}

// Mixed form: UBM-based centring o = sum_m w_m mu_m and scaling
// s = 1 / sqrt(sum_m w_m (var_m + (mu_m - o)^2)) (the UBM's global mean and
// deviation), and the split model for them.  Returns false when the model's
// coefficients overflow the 16-bit type (the server then runs FP32 batches).
int setup_mixed(gmmhip_server *s, bool *ok) {
    gmmhip_ctx *b = s->base;
    const int M = s->d.M, D = s->d.D;
    s->h_zoff.assign(D, 0.0f);
    s->h_zscale.assign(D, 1.0f);
    double wsum = 0.0;
    for (int m = 0; m < M; m++) wsum += b->m_w[m];
    if (!(wsum > 0.0)) wsum = 1.0;
    for (int d = 0; d < D; d++) {
        double o = 0.0;
        for (int m = 0; m < M; m++) o += b->m_w[m] * (double)b->m_mu[(size_t)m * D + d];
        o /= wsum;
        double v = 0.0;
        for (int m = 0; m < M; m++) {
            const double dm = (double)b->m_mu[(size_t)m * D + d] - o;
            v += b->m_w[m] * ((double)b->m_var[(size_t)m * D + d] + dm * dm);
        }
        v /= wsum;
        s->h_zoff[d] = (float)o;
        if (std::isfinite(v) && v > 0.0) s->h_zscale[d] = (float)(1.0 / std::sqrt(v));
    }
    SRV_HIPCHK(hipMalloc((void **)&s->d_zoff, D * 4));
    SRV_HIPCHK(hipMalloc((void **)&s->d_zscale, D * 4));
    SRV_HIPCHK(hipMalloc(&s->d_P16, (size_t)((M + 31) / 32) * s->d.KPm * 64 * 2));
    SRV_HIPCHK(hipMemcpy(s->d_zoff, s->h_zoff.data(), D * 4, hipMemcpyHostToDevice));
    SRV_HIPCHK(hipMemcpy(s->d_zscale, s->h_zscale.data(), D * 4, hipMemcpyHostToDevice));
    bool overflow = false;
    const int rc = pack_split_model(b, s->prec, s->d.KPm, s->d_zoff, s->d_zscale,
                                    static_cast<uint16_t *>(s->d_P16), &overflow);
    if (rc) return rc;
    *ok = !overflow;
    if (overflow) return GMMHIP_OK;
    // frames with |z| above this run in FP32 (prep_frames flags > 240)
    s->zlimit = s->prec == GMMHIP_PREC_FP16X3 ? 240.0f : INFINITY;
    return GMMHIP_OK;
}

// True when every finite value of the request maps inside the FP16 range
// with a margin (the device flags |z| > 240).
bool fits_fp16(const gmmhip_server *s, const gmmhip_request *r) {
    const int D = s->d.D;
    const float *o = s->h_zoff.data(), *sc = s->h_zscale.data();
    for (uint64_t t = 0; t < r->n_frames; t++) {
        const float *x = r->frames + t * D;
        bool ok = true;
        for (int d = 0; d < D; d++) {
            const float z = (x[d] - o[d]) * sc[d];
            // non-finite values pass: the kernels flag those frames
            ok &= !(std::isfinite(x[d]) && std::fabs(z) > 239.0f);
        }
        if (!ok) return false;
    }
    return true;
}

}  // namespace

extern "C" int gmmhip_server_create(int device, const gmmhip_server_cfg *cfg, int M, int D,
                                    const float *w, const float *mu, const float *var,
                                    gmmhip_server_t *out) {
    if (!out) return GMMHIP_ERR_ARG;
    *out = nullptr;
    gmmhip_server_cfg c{};
    if (cfg) c = *cfg;
    else gmmhip_default_opts(&c.opts);
    const int prec = c.opts.precision;
    if (prec != GMMHIP_PREC_FP32 && prec != GMMHIP_PREC_BF16X3 && prec != GMMHIP_PREC_FP16X3)
        return GMMHIP_ERR_ARG;
    if (c.slots <= 0) c.slots = 2;
    if (c.slots > 8) return GMMHIP_ERR_ARG;
    if (c.slot_bytes == 0) c.slot_bytes = 1ull << 30;
    if (c.max_batch <= 0) c.max_batch = 1024;
    if (c.chunk_frames < 0 || c.tau < 0) return GMMHIP_ERR_ARG;
    if (M > MLLR_MAX_COMPONENTS) return GMMHIP_ERR_ARG;  // MLLR requests
    gmmhip_server *s = new gmmhip_server();
    s->cfg = c;
    s->dev = device;
    s->cf = c.chunk_frames > 0 ? (uint64_t)(c.chunk_frames + 63) / 64 * 64 : 512;
    int rc = gmmhip_create(device, &c.opts, &s->base);
    if (rc) {
        delete s;
        return rc;
    }
    if ((rc = gmmhip_set_model(s->base, M, D, w, mu, var))) {
        free_server(s);
        return rc;
    }
    gmmhip_ctx *b = s->base;
    DevGuard g(device);
    if ((rc = model_operands(b, true, &s->mo))) {
        free_server(s);
        return rc;
    }
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
        ncu > 0)
        s->num_cus = ncu;
    s->d.M = M;
    s->d.D = D;
    s->d.E = D + 1;
    s->d.NP = (D + 1) * (D + 2) / 2;
    s->d.KP = b->KP;
    s->d.Mp = b->Mt * 16;
    s->d.KPm = kp_mixed_for_dim(D);
    s->d.Mpm = (M + 31) / 32 * 32;
    if (prec != GMMHIP_PREC_FP32) {
        s->prec = prec;
        bool ok = false;
        if ((rc = setup_mixed(s, &ok))) {
            free_server(s);
            return rc;
        }
        if (!ok) {
            s->prec = GMMHIP_PREC_FP32;
            s->err = "model coefficients overflow the 16-bit type: serving in FP32";
        }
    }
    s->slots.resize(c.slots);
    for (int i = 0; i < c.slots; i++) {
        auto &sl = s->slots[i];
        sl.id = i;
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreate(&sl.ev0) != hipSuccess || hipEventCreate(&sl.ev1) != hipSuccess) {
            free_server(s);
            return GMMHIP_ERR_HIP;
        }
        if (hipMalloc((void **)&sl.arena, c.slot_bytes) != hipSuccess) {
            free_server(s);
            return GMMHIP_ERR_OOM;
        }
        // one pinned staging buffer of the slot budget: every admitted
        // batch's IN and OUT regions fit in it, so it never regrows on the
        // dispatch path (a regrowth frees pinned memory, which waits for the
        // device, and pins anew)
        if (hipHostMalloc((void **)&sl.h_stage, c.slot_bytes, hipHostMallocDefault) != hipSuccess) {
            sl.h_stage = nullptr;
            free_server(s);
            return GMMHIP_ERR_OOM;
        }
        sl.h_stage_cap = c.slot_bytes;
    }
    s->t_reset = Clock::now();
    for (int i = 0; i < c.slots; i++) s->slots[i].completer = std::thread(completer_main, s, i);
    s->dispatcher = std::thread(dispatcher_main, s);
    *out = s;
    return GMMHIP_OK;
}

extern "C" int gmmhip_server_submit(gmmhip_server_t s, gmmhip_request *r) {
    if (!s || !r) return GMMHIP_ERR_ARG;
    if (r->kind < GMMHIP_REQ_SCORE || r->kind > GMMHIP_REQ_MLLR) return GMMHIP_ERR_ARG;
    if (!r->frames || r->n_frames == 0) return GMMHIP_ERR_ARG;
    if (r->kind == GMMHIP_REQ_MAP && !r->means) return GMMHIP_ERR_ARG;
    if (r->kind == GMMHIP_REQ_MLLR && !r->mllr_W) return GMMHIP_ERR_ARG;
    // the form is decided by the request's own frames
    bool mixed = s->prec != GMMHIP_PREC_FP32;
    if (mixed && s->prec == GMMHIP_PREC_FP16X3) mixed = fits_fp16(s, r);
    std::lock_guard<std::mutex> g(s->mu);
    if (s->closing) return GMMHIP_ERR_CLOSED;
    r->ll_sum = 0.0;
    r->latency_ms = 0.0;
    __atomic_store_n(&r->status, GMMHIP_REQ_PENDING, __ATOMIC_RELEASE);
    (mixed ? s->q_mixed : s->q_fp32).push_back(Pending{r, Clock::now()});
    if (s->prec != GMMHIP_PREC_FP32 && !mixed) s->fp32_fallbacks++;
    s->submitted++;
    s->cv_dispatch.notify_one();
    return GMMHIP_OK;
}

extern "C" int gmmhip_server_wait(gmmhip_server_t s, gmmhip_request *r, double timeout_ms) {
    if (!s || !r) return GMMHIP_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto done = [&] { return __atomic_load_n(&r->status, __ATOMIC_ACQUIRE) != GMMHIP_REQ_PENDING; };
    if (timeout_ms < 0) {
        s->cv_done.wait(lk, done);
    } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms),
                                    done)) {
        return GMMHIP_ERR_TIMEOUT;
    }
    return r->status;
}

extern "C" int gmmhip_server_drain(gmmhip_server_t s) {
    if (!s) return GMMHIP_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->completed + s->failed == s->submitted; });
    return GMMHIP_OK;
}

extern "C" int gmmhip_server_stats_get(gmmhip_server_t s, gmmhip_server_stats *st, int reset) {
    if (!s || !st) return GMMHIP_ERR_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    *st = gmmhip_server_stats{};
    st->submitted = s->submitted;
    st->completed = s->completed;
    st->failed = s->failed;
    st->batches = s->batches;
    st->frames = s->frames;
    st->fp32_fallbacks = s->fp32_fallbacks;
    std::vector<double> v = s->lat;
    std::sort(v.begin(), v.end());
    auto pct = [&v](double q) {
        if (v.empty()) return 0.0;
        size_t i = (size_t)std::ceil(q * v.size());
        i = i ? i - 1 : 0;
        return v[std::min(i, v.size() - 1)];
    };
    st->p50_ms = pct(0.50);
    st->p90_ms = pct(0.90);
    st->p99_ms = pct(0.99);
    st->p999_ms = pct(0.999);
    st->max_ms = v.empty() ? 0.0 : v.back();
    double sum = 0.0;
    for (double x : v) sum += x;
    st->mean_ms = v.empty() ? 0.0 : sum / v.size();
    st->mean_batch_requests = s->batches ? (double)s->batch_requests / s->batches : 0.0;
    st->mean_batch_gpu_ms = s->batches ? s->batch_gpu_ms / s->batches : 0.0;
    st->wall_s = std::chrono::duration<double>(Clock::now() - s->t_reset).count();
    st->precision = s->prec;
    if (reset) {
        s->lat.clear();
        s->submitted -= s->completed + s->failed;
        s->completed = s->failed = s->batches = s->frames = s->batch_requests = 0;
        s->fp32_fallbacks = 0;
        s->batch_gpu_ms = 0.0;
        s->t_reset = Clock::now();
    }
    return GMMHIP_OK;
}

extern "C" const char *gmmhip_server_last_error(gmmhip_server_t s) {
    if (!s) return "null server";
    std::lock_guard<std::mutex> g(s->mu);
    return s->err.c_str();
}

extern "C" void gmmhip_server_destroy(gmmhip_server_t s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->closing = true;
        s->cv_dispatch.notify_all();
    }
    if (s->dispatcher.joinable()) s->dispatcher.join();
    for (auto &sl : s->slots)
        if (sl.completer.joinable()) sl.completer.join();
    free_server(s);
}
