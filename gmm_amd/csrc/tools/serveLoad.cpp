// serveLoad: open-loop load generator for the serving loop (gmmhip_server_*).
//
// A seeded UBM (weights ~ Dirichlet(0.5), means ~ N(0, 2^2), variances
// ~ U[0.3, 3]) is made resident; a pool of utterances is drawn, each from its
// own perturbed mixture (means + N(0, 0.3^2)).  For each rate, requests
// arrive as a Poisson process (exponential gaps), kinds 60% score / 30% MAP /
// 10% MLLR by default, lengths ~ lognormal(ln median, sigma) clipped to
// [min, max] frames, frames taken from a random pool utterance at a random
// offset.  Latency is measured from the scheduled arrival time to
// completion, so a submitter that falls behind counts against the server.
// The first requests after the warm-up are run again one at a time and
// their results compared bit for bit with the ones they got in the mix.
// One JSON line per rate on stdout.
#include "../../../include/gmmhip.h"
#include "cli_args.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

using Clock = std::chrono::steady_clock;

namespace {

std::vector<double> parse_list(const std::string &s) {
    std::vector<double> v;
    std::stringstream ss(s);
    std::string t;
    while (std::getline(ss, t, ','))
        if (!t.empty()) v.push_back(std::atof(t.c_str()));
    return v;
}

// Words-at-a-time mixing hash of result bytes (for the alone-vs-mix check).
uint64_t hash_bytes(const void *p, size_t n, uint64_t h) {
    const unsigned char *b = (const unsigned char *)p;
    size_t i = 0;
    for (; i + 8 <= n; i += 8) {
        uint64_t w;
        std::memcpy(&w, b + i, 8);
        h = (h ^ w) * 0x100000001b3ull;
        h ^= h >> 29;
    }
    for (; i < n; i++) h = (h ^ b[i]) * 0x100000001b3ull;
    return h;
}

struct Model {
    int M, D;
    std::vector<float> w, mu, var;
};

Model make_ubm(int M, int D, uint64_t seed) {
    Model m{M, D, std::vector<float>(M), std::vector<float>((size_t)M * D),
            std::vector<float>((size_t)M * D)};
    std::mt19937_64 g(seed);
    std::gamma_distribution<double> gam(0.5, 1.0);
    std::normal_distribution<float> nrm(0.0f, 2.0f);
    std::uniform_real_distribution<float> uni(0.3f, 3.0f);
    std::vector<double> wd(M);
    double s = 0.0;
    for (int k = 0; k < M; k++) s += (wd[k] = std::max(gam(g), 1e-300));
    for (int k = 0; k < M; k++) m.w[k] = (float)(wd[k] / s);
    for (auto &v : m.mu) v = nrm(g);
    for (auto &v : m.var) v = uni(g);
    return m;
}

// One pool utterance of L frames from the UBM with means shifted by N(0, 0.3^2).
void draw_utterance(const Model &m, uint64_t seed, size_t L, float *out) {
    std::mt19937_64 g(seed);
    std::normal_distribution<float> sh(0.0f, 0.3f), z(0.0f, 1.0f);
    std::vector<float> mu(m.mu);
    for (auto &v : mu) v += sh(g);
    std::vector<double> cdf(m.M);
    double c = 0.0;
    for (int k = 0; k < m.M; k++) cdf[k] = (c += m.w[k]);
    std::uniform_real_distribution<double> u01(0.0, c);
    for (size_t t = 0; t < L; t++) {
        const int k = (int)(std::lower_bound(cdf.begin(), cdf.end(), u01(g)) - cdf.begin());
        const int kk = std::min(k, m.M - 1);
        for (int d = 0; d < m.D; d++)
            out[t * m.D + d] = mu[(size_t)kk * m.D + d] + std::sqrt(m.var[(size_t)kk * m.D + d]) * z(g);
    }
}

struct Planned {
    double t;  // arrival, seconds from the start
    int kind;
    uint32_t len, pool;
    uint64_t off;
};

struct Slot {
    gmmhip_request r;
    std::vector<float> means;
    std::vector<double> W;
    long idx = -1;
};

double pct(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    size_t i = (size_t)std::ceil(q * v.size());
    i = i ? i - 1 : 0;
    return v[std::min(i, v.size() - 1)];
}

uint64_t result_hash(const gmmhip_request &r, int M, int D) {
    uint64_t h = 1469598103934665603ull;
    h = hash_bytes(&r.ll_sum, 8, h);
    const int st = r.status;
    h = hash_bytes(&st, 4, h);
    if (r.kind == GMMHIP_REQ_MAP) h = hash_bytes(r.means, (size_t)M * D * 4, h);
    if (r.kind == GMMHIP_REQ_MLLR) h = hash_bytes(r.mllr_W, (size_t)D * (D + 1) * 8, h);
    return h;
}

}  // namespace

int main(int argc, char **argv) {
    gmm::CliArgs a;
    a.option("M", 0, false, "2048");
    a.option("D", 0, false, "60");
    a.option("rates", 0, false, "2000,5000,10000");
    a.option("seconds", 0, false, "5");
    a.option("warmup", 0, false, "1");
    a.option("seed", 0, false, "20261208");
    a.option("median", 0, false, "300");
    a.option("sigma", 0, false, "1.0");
    a.option("min-frames", 0, false, "20");
    a.option("max-frames", 0, false, "6000");
    a.option("mix", 0, false, "0.6,0.3,0.1");
    a.option("slots", 0, false, "2");
    a.option("slot-mb", 0, false, "1024");
    a.option("chunk", 0, false, "512");
    a.option("max-batch", 0, false, "1024");
    a.option("tau", 0, false, "16");
    a.option("prior", 0, true, "0");
    a.option("precision", 0, false, "fp32");
    a.option("pool", 0, false, "64");
    a.option("ring", 0, false, "2048");
    a.option("verify", 0, false, "64");
    a.option("device", 0, false, "0");
    try {
        a.parse(argc, argv);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "serveLoad: %s\n", e.what());
        return 2;
    }
    const int M = a.i("M"), D = a.i("D"), E = D + 1;
    const uint64_t seed = (uint64_t)std::atoll(a.str("seed").c_str());
    const std::vector<double> rates = parse_list(a.str("rates"));
    const double seconds = a.f("seconds"), warm = a.f("warmup");
    const std::vector<double> mix = parse_list(a.str("mix"));
    const uint32_t minf = (uint32_t)a.i("min-frames"), maxf = (uint32_t)a.i("max-frames");
    const int npool = a.i("pool"), ring = a.i("ring"), nverify = a.i("verify");
    if (mix.size() != 3 || M <= 0 || D <= 0 || npool <= 0 || ring <= 0 || maxf < minf) {
        std::fprintf(stderr, "serveLoad: bad arguments\n");
        return 2;
    }
    const std::string prec = a.str("precision");

    // model and utterance pool
    const Model ubm = make_ubm(M, D, seed);
    std::vector<float> pool((size_t)npool * maxf * D);
    {
        std::vector<std::thread> th;
        const int nt = (int)std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency()));
        for (int t = 0; t < nt; t++)
            th.emplace_back([&, t] {
                for (int p = t; p < npool; p += nt)
                    draw_utterance(ubm, seed + 1000 + p, maxf, &pool[(size_t)p * maxf * D]);
            });
        for (auto &x : th) x.join();
    }
    std::fprintf(stderr, "serveLoad: UBM M=%d D=%d, pool %d x %u frames\n", M, D, npool, maxf);

    gmmhip_server_cfg cfg{};
    gmmhip_default_opts(&cfg.opts);
    cfg.opts.precision = prec == "fp16x3" ? GMMHIP_PREC_FP16X3
                         : prec == "bf16x3" ? GMMHIP_PREC_BF16X3
                                            : GMMHIP_PREC_FP32;
    cfg.slots = a.i("slots");
    cfg.slot_bytes = (uint64_t)a.i("slot-mb") << 20;
    cfg.max_batch = a.i("max-batch");
    cfg.chunk_frames = a.i("chunk");
    cfg.tau = a.i("tau");
    cfg.mllr_prior = a.b("prior") ? 1 : 0;
    gmmhip_server_t srv = nullptr;
    int rc = gmmhip_server_create(a.i("device"), &cfg, M, D, ubm.w.data(), ubm.mu.data(),
                                  ubm.var.data(), &srv);
    if (rc) {
        std::fprintf(stderr, "serveLoad: gmmhip_server_create failed (%d)\n", rc);
        return 1;
    }

    std::vector<Slot> slots(ring);
    for (auto &s : slots) {
        s.means.assign((size_t)M * D, 0.0f);
        s.W.assign((size_t)D * E, 0.0);
    }

    int exit_code = 0;
    for (size_t ri = 0; ri < rates.size(); ri++) {
        const double rate = rates[ri];
        std::mt19937_64 g(seed * 7919 + ri);
        std::exponential_distribution<double> gap(rate);
        std::lognormal_distribution<double> lenD(std::log(a.f("median")), a.f("sigma"));
        std::uniform_real_distribution<double> u01(0.0, 1.0);
        std::vector<Planned> plan;
        for (double t = gap(g); t < warm + seconds; t += gap(g)) {
            Planned p;
            p.t = t;
            const double k = u01(g);
            p.kind = k < mix[0] ? GMMHIP_REQ_SCORE
                     : k < mix[0] + mix[1] ? GMMHIP_REQ_MAP
                                           : GMMHIP_REQ_MLLR;
            p.len = (uint32_t)std::min<double>(maxf, std::max<double>(minf, std::round(lenD(g))));
            p.pool = (uint32_t)(g() % npool);
            p.off = (maxf > p.len) ? g() % (maxf - p.len + 1) : 0;
            plan.push_back(p);
        }
        const long n = (long)plan.size();
        std::vector<double> lat(n, -1.0), lag(n, 0.0);
        std::vector<int> status(n, 0);
        std::vector<uint64_t> hmix(n, 0);
        long first_measured = 0;
        while (first_measured < n && plan[first_measured].t < warm) first_measured++;
        const long verify_end = std::min<long>(n, first_measured + nverify);
        uint64_t frames_measured = 0;

        auto collect = [&](Slot &s) {
            if (s.idx < 0) return;
            gmmhip_server_wait(srv, &s.r, -1.0);
            lat[s.idx] = lag[s.idx] + s.r.latency_ms;
            status[s.idx] = s.r.status;
            if (s.idx >= first_measured && s.idx < verify_end) hmix[s.idx] = result_hash(s.r, M, D);
            s.idx = -1;
        };

        gmmhip_server_stats st0;
        gmmhip_server_stats_get(srv, &st0, 1);
        const Clock::time_point t0 = Clock::now() + std::chrono::milliseconds(20);
        Clock::time_point t_meas = t0;
        for (long i = 0; i < n; i++) {
            Slot &s = slots[i % ring];
            collect(s);
            const Planned &p = plan[i];
            std::memset(&s.r, 0, sizeof(s.r));
            s.r.kind = p.kind;
            s.r.frames = &pool[((size_t)p.pool * maxf + p.off) * D];
            s.r.n_frames = p.len;
            if (p.kind == GMMHIP_REQ_MAP) s.r.means = s.means.data();
            if (p.kind == GMMHIP_REQ_MLLR) s.r.mllr_W = s.W.data();
            s.idx = i;
            const Clock::time_point due =
                t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(p.t));
            if (i == first_measured) {
                // statistics window starts at the first measured arrival
                gmmhip_server_stats_get(srv, &st0, 1);
                t_meas = due;
            }
            Clock::time_point now = Clock::now();
            if (due - now > std::chrono::microseconds(300))
                std::this_thread::sleep_until(due - std::chrono::microseconds(200));
            while ((now = Clock::now()) < due) {
            }
            lag[i] = std::chrono::duration<double, std::milli>(now - due).count();
            if (i >= first_measured) frames_measured += p.len;
            if ((rc = gmmhip_server_submit(srv, &s.r))) {
                std::fprintf(stderr, "serveLoad: submit failed (%d)\n", rc);
                return 1;
            }
        }
        for (auto &s : slots) collect(s);
        const double elapsed = std::chrono::duration<double>(Clock::now() - t_meas).count();
        gmmhip_server_stats st;
        gmmhip_server_stats_get(srv, &st, 1);

        // the same requests again, one at a time
        long identical = 0, checked = 0;
        for (long i = first_measured; i < verify_end; i++) {
            Slot &s = slots[0];
            const Planned &p = plan[i];
            std::memset(&s.r, 0, sizeof(s.r));
            s.r.kind = p.kind;
            s.r.frames = &pool[((size_t)p.pool * maxf + p.off) * D];
            s.r.n_frames = p.len;
            if (p.kind == GMMHIP_REQ_MAP) s.r.means = s.means.data();
            if (p.kind == GMMHIP_REQ_MLLR) s.r.mllr_W = s.W.data();
            gmmhip_server_submit(srv, &s.r);
            gmmhip_server_wait(srv, &s.r, -1.0);
            checked++;
            if (result_hash(s.r, M, D) == hmix[i]) identical++;
        }
        gmmhip_server_stats_get(srv, &st0, 1);

        std::vector<double> all, bykind[3];
        long failed = 0, mllr_under = 0;
        double lagmax = 0.0;
        for (long i = first_measured; i < n; i++) {
            if (status[i] != GMMHIP_OK && status[i] != GMMHIP_ERR_MLLR) {
                failed++;
                continue;
            }
            if (status[i] == GMMHIP_ERR_MLLR) mllr_under++;
            all.push_back(lat[i]);
            bykind[plan[i].kind].push_back(lat[i]);
            lagmax = std::max(lagmax, lag[i]);
        }
        const long measured = n - first_measured;
        std::printf(
            "{\"rate\": %.0f, \"precision\": \"%s\", \"M\": %d, \"D\": %d, \"seconds\": %.2f, "
            "\"requests\": %ld, \"failed\": %ld, \"mllr_underdetermined\": %ld, "
            "\"throughput_rps\": %.1f, \"frames_per_s\": %.4g, \"frame_comp_per_s\": %.4g, "
            "\"p50_ms\": %.3f, \"p90_ms\": %.3f, \"p99_ms\": %.3f, \"p999_ms\": %.3f, "
            "\"max_ms\": %.3f, "
            "\"score_p99_ms\": %.3f, \"map_p50_ms\": %.3f, \"map_p99_ms\": %.3f, "
            "\"mllr_p99_ms\": %.3f, \"submit_lag_max_ms\": %.3f, "
            "\"batches\": %llu, \"mean_batch_requests\": %.2f, \"mean_batch_gpu_ms\": %.3f, "
            "\"verify_checked\": %ld, \"verify_identical\": %ld}\n",
            rate, prec.c_str(), M, D, elapsed, measured, failed, mllr_under,
            (measured - failed) / elapsed, frames_measured / elapsed,
            (double)frames_measured * M / elapsed, pct(all, 0.5), pct(all, 0.9), pct(all, 0.99),
            pct(all, 0.999), all.empty() ? 0.0 : *std::max_element(all.begin(), all.end()),
            pct(bykind[0], 0.99), pct(bykind[1], 0.5), pct(bykind[1], 0.99), pct(bykind[2], 0.99),
            lagmax, (unsigned long long)st.batches, st.mean_batch_requests, st.mean_batch_gpu_ms,
            checked, identical);
        std::fflush(stdout);
        if (failed || identical != checked) exit_code = 3;
    }
    gmmhip_server_destroy(srv);
    return exit_code;
}
